"""The #pragma gtap lowering tool (host logic, no GPU): the examples lower to the expected
state machines, the built libraries export their entry points, and unsupported forms are
rejected with an error rather than miscompiled."""
import ctypes
import glob
import os

import pytest

from gtap_amd import build as gbuild
from gtap_amd import lower

EX = os.path.join(os.path.dirname(lower.__file__), "examples")


def read(name):
    with open(os.path.join(EX, name)) as f:
        return f.read()


def test_fib_lowering_structure():
    out = lower.lower(read("fib.gtap.cu"), "fib.gtap.cu")
    assert "constexpr uint32_t F_fib = 1;" in out
    assert "struct fib__frame" in out and "long long a;" in out and "long long b;" in out
    # two spawns naming the result slots of a and b, one taskwait with resume state 1
    assert out.count("__gtap_spawn_task(lane, F_fib") == 2
    assert "offsetof(fib__frame, a)" in out and "offsetof(fib__frame, b)" in out
    assert "__gtap_set_state_for_join(lane, 1u" in out and "case 1: fib__load(lane, fr);" in out
    # the result slots are loaded after the join and not saved before it
    save = out[out.index("void fib__save"):out.index("void fib__load")]
    assert "frame_st" not in save
    assert 'extern "C" int gtap_entry_fib_kernel(int n, long long* result, gtap_run_stats* stats)' in out
    assert "long long fib__seq(int n)" in out


def test_nqueens_array_result_slots():
    out = lower.lower(read("nqueens.gtap.cu"), "nqueens.gtap.cu")
    assert "offsetof(nqueens__frame, counts) + 8 * (fr.k++)" in out
    save = out[out.index("void nqueens__save"):out.index("void nqueens__load")]
    assert "counts" not in save  # children write them after the suspension
    assert "fr.k" in save and "fr.avail" in save
    load = out[out.index("void nqueens__load"):]
    assert "fr.counts[i] = lane.template frame_ld<long long>" in load


def test_void_tasks_and_entry():
    out = lower.lower(read("msort.gtap.cu"), "msort.gtap.cu")
    assert out.count("offsetof(msort__frame, __discard)") == 2  # joined children without a result
    assert "__gtap_finish_task(lane); return;" in out
    assert 'extern "C" int gtap_entry_msort_kernel(unsigned int * a, unsigned int * tmp, long n, gtap_run_stats* stats)' in out
    assert "__device__ void merge_into(" in out  # device helpers are carried over


def test_lowering_is_deterministic():
    src = read("nqueens.gtap.cu")
    assert lower.lower(src, "a.cu", "x") == lower.lower(src, "a.cu", "x")


BAD = {
    "no task call": """
#pragma gtap function
__device__ int f(int n) {
    int a;
    #pragma gtap task
    a = n + 1;
    #pragma gtap taskwait
    return a;
}
""",
    "result slot size": """
#pragma gtap function
__device__ long long f(int n) {
    if (n < 1) return 0;
    int a;
    #pragma gtap task
    a = f(n - 1);
    #pragma gtap taskwait
    return a;
}
""",
    "array initialiser": """
#pragma gtap function
__device__ int f(int n) {
    int v[4] = {1, 2, 3, 4};
    return v[n & 3];
}
""",
    "parse error": """
#pragma gtap function
__device__ int f(int n) {
    return n +;
}
""",
    "no function": "__device__ int g(int x) { return x; }\n",
}


@pytest.mark.parametrize("case", sorted(BAD))
def test_rejected_forms(case):
    with pytest.raises(lower.LoweringError):
        lower.lower(BAD[case], "bad.cu")


def test_built_libraries_export_entries():
    for src in sorted(glob.glob(os.path.join(EX, "*.gtap.cu"))):
        name = os.path.basename(src).split(".")[0]
        path = gbuild.lowered_lib(name)
        assert os.path.exists(path), f"{path}: run python -m gtap_amd.build"
        text = lower.lower(open(src).read(), src, name)
        entries = [l.split("(")[0].split()[-1] for l in text.splitlines() if l.startswith('extern "C" int gtap_entry_')]
        L = ctypes.CDLL(os.path.join(os.path.dirname(path), "libgtap_amd.so"))
        L2 = ctypes.CDLL(path)
        for e in entries + [f"gtap_lowered_{name}_record_bytes"]:
            assert hasattr(L2, e), (path, e)
        del L, L2
