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


def test_int_sum_of_children_joins_additively():
    """A task whose 32-bit children's results are only summed after one taskwait (the
    reference's fib.cu) is lowered to the additive join: no result slots, the sum read from
    the join word."""
    out = lower.lower(read("fib32.gtap.cu"), "fib32.gtap.cu")
    assert out.count("__gtap_spawn_task_sum(lane, F_fib") == 2
    assert "__gtap_spawn_task(lane" not in out
    assert "__gtap_set_state_for_join_sum(lane, 1u" in out
    assert "__gtap_return(lane, (int)(__gtap_get_join_sum(lane)))" in out
    load = out[out.index("void fib__load"):out.index("constexpr uint32_t kMaxOf")]
    assert "frame_ld" not in load  # nothing to read back after the join
    assert "__gtap_return(lane, (int)(fr.n))" in out  # the early return is unchanged


SUM_SHAPE = """#include "gtap_thread.cuh"
#pragma gtap function
__device__ RET f(int n) {
    if (n < 2) return n;
    DECL;
    BODY
}
__global__ void k(int n) {
    #pragma gtap entry
    f(n);
}
"""

KEEPS_SLOTS = {
    "64-bit result": ("long long", "long long a, b",
                      "#pragma gtap task\n a = f(n - 1);\n #pragma gtap task\n b = f(n - 2);\n"
                      " #pragma gtap taskwait\n return a + b;"),
    "task under a condition": ("int", "int a = 0, b = 0",
                               "#pragma gtap task\n a = f(n - 1);\n if (n > 4) {\n #pragma gtap task\n b = f(n - 2);\n }\n"
                               " #pragma gtap taskwait\n return a + b;"),
    "other use of a result": ("int", "int a, b",
                              "#pragma gtap task\n a = f(n - 1);\n #pragma gtap task\n b = f(n - 2);\n"
                              " #pragma gtap taskwait\n if (a > 3) return a;\n return a + b;"),
    "not a plain sum": ("int", "int a, b",
                        "#pragma gtap task\n a = f(n - 1);\n #pragma gtap task\n b = f(n - 2);\n"
                        " #pragma gtap taskwait\n return a + b + 1;"),
    "product": ("int", "int a, b",
                "#pragma gtap task\n a = f(n - 1);\n #pragma gtap task\n b = f(n - 2);\n"
                " #pragma gtap taskwait\n return a * b;"),
    "two taskwaits": ("int", "int a, b",
                      "#pragma gtap task\n a = f(n - 1);\n #pragma gtap taskwait\n"
                      " #pragma gtap task\n b = f(n - 2);\n #pragma gtap taskwait\n return a + b;"),
    "task in a loop": ("int", "int a = 0, i",
                       "for (i = 0; i < 2; ++i) {\n #pragma gtap task\n a = f(n - 1 - i);\n }\n"
                       " #pragma gtap taskwait\n return a;"),
}


@pytest.mark.parametrize("case", sorted(KEEPS_SLOTS))
def test_other_shapes_keep_result_slots(case):
    ret, decl, body = KEEPS_SLOTS[case]
    src = SUM_SHAPE.replace("RET", ret).replace("DECL", decl).replace("BODY", body.replace("\\n", "\n"))
    out = lower.lower(src, "shape.cu")
    assert "__gtap_spawn_task_sum" not in out and "__gtap_get_join_sum" not in out
    assert "__gtap_spawn_task(lane, F_f" in out


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


PLAIN_MSORT = """
#include "gtap_thread.cuh"
__device__ int* g_data;
__device__ int* g_buf;
__device__ void merge(int* a, int* buf, int l, int m, int r) {
    int i = l, j = m + 1, k = l;
    while (i <= m && j <= r) {
        if (a[i] <= a[j]) buf[k++] = a[i++];
        else buf[k++] = a[j++];
    }
    while (i <= m) buf[k++] = a[i++];
    for (int t = l; t <= r; ++t) a[t] = buf[t];
}
#pragma gtap function
__device__ void msort(int l, int r) {
    if (l >= r) return;
    int m = l + (r - l) / 2;
    #pragma gtap task
    msort(l, m);
    #pragma gtap task
    msort(m + 1, r);
    #pragma gtap taskwait
    merge(g_data, g_buf, l, m, r);
}
__global__ void k(int n) {
    g_buf[0] = load_L2(&g_data[0]);
    #pragma gtap entry
    msort(0, n - 1);
}
"""


def test_plain_accesses_become_coherent():
    """Task code reads and writes shared data with plain accesses (the reference's idiom,
    examples/mergesort/mergesort.cu:15-24): every one becomes a coherent load or store, the
    address of an element stays an address, and the kernel's own statements are rewritten too."""
    out = lower.lower(PLAIN_MSORT, "m.cu")
    merge = out[out.index("__device__ void merge("):out.index("__device__ void msort__seq(int l, int r) {")]
    assert "::gtap::cld(&(a[i])) <= ::gtap::cld(&(a[j]))" in merge
    assert "::gtap::coh(buf[k++]) = ::gtap::cld(&(a[i++]))" in merge
    assert "::gtap::coh(a[t]) = ::gtap::cld(&(buf[t]))" in merge
    kern = out[out.index("void k(int n)"):]
    assert "::gtap::coh(g_buf[0]) = load_L2(&g_data[0]);" in kern
    assert "a[i]" not in merge.replace("&(a[i])", "").replace("(a[i++])", "")


def test_coherent_rewrite_can_be_turned_off():
    out = lower.lower(PLAIN_MSORT, "m.cu", coherent=False)
    assert "::gtap::cld(&(" not in out and "::gtap::coh(" not in out
    assert "if (a[i] <= a[j]) buf[k++] = a[i++];" in out


def test_device_variables_entry_kernel_and_symbols():
    out = lower.lower(read("fib_epaq.gtap.cu"), "fib_epaq.gtap.cu")
    assert "__device__ long long d_result;" in out and "__device__ __constant__ int d_cutoff;" in out
    assert "static constexpr int kQueues = 3;" in out  # #define GTAP_NUM_QUEUES 3 in the source
    assert "__global__ __launch_bounds__(GTAP_BLOCK_SIZE) void fib_kernel(int n)" in out
    assert "d_result = ::gtap::result_as<long long>(__gtap_bits);" in out
    assert "(uint32_t)((fr.n - 1) < d_cutoff ? 1 : 0)" in out  # queue(expr) over a parameter
    assert "__gtap_set_state_for_join(lane, 1u, (uint32_t)(2))" in out
    assert 'if (strcmp(name, "d_cutoff") == 0)' in out
    assert 'extern "C" int gtap_launch_fib_kernel(uint32_t grid, uint32_t block, int n)' in out
    # every spawn is joined, so the root's end is the last task's end and the job count ends it
    assert "constexpr uint32_t kTermMode = GTAP_TERM_JOBS;" in out


def test_switches_from_the_command_line():
    src = read("fib.gtap.cu")
    out = lower.lower(src, "fib.gtap.cu", defines={"GTAP_NUM_QUEUES": "2", "GTAP_TERMINATE_ON_FIRST_TASK_FINISH": ""})
    assert "#define GTAP_NUM_QUEUES 2" in out and "static constexpr int kQueues = 2;" in out
    assert "constexpr uint32_t kTermMode = GTAP_TERM_JOBS;" in out
    with pytest.raises(lower.LoweringError):
        lower.lower(src, "fib.gtap.cu", defines={"GTAP_BLOCK_SIZE": "32"})  # 32-lane warps do not exist here
    assert "amdgpu_waves_per_eu" not in out and "kWavesPerSimd" not in out
    out = lower.lower(src, "fib.gtap.cu", defines={"GTAP_WAVES_PER_SIMD": "4"})
    assert "static constexpr int kWavesPerSimd = 4;" in out
    assert "__attribute__((amdgpu_waves_per_eu(4))) void fib_kernel(" in out
    for bad in ("0", "9", "four"):
        with pytest.raises(lower.LoweringError):
            lower.lower(src, "fib.gtap.cu", defines={"GTAP_WAVES_PER_SIMD": bad})


def test_detached_spawns_without_taskwait():
    """A spawn that no taskwait follows is detached (examples/nq/nq.cu:45-49): the lowering
    uses the detached spawn, and a program without any taskwait is built without the
    suspension staging."""
    out = lower.lower(lower_src_nq(), "nq.cu")
    assert "__gtap_spawn_task_detached(lane, F_nq" in out and "__gtap_spawn_task(lane" not in out
    assert "static constexpr bool kNoTaskwait = true;" in out
    # detached tasks may outlive the root: the launch ends when no task is left anywhere
    assert "constexpr uint32_t kTermMode = GTAP_TERM_QUIESCE;" in out


def lower_src_nq():
    return """
__device__ int d_answer;
#pragma gtap function
__device__ void nq(int n, int row, unsigned cols) {
    if (row == n) { atomicAdd(&d_answer, 1); return; }
    for (int c = 0; c < n; ++c) {
        if (cols & (1u << c)) continue;
        #pragma gtap task
        nq(n, row + 1, cols | (1u << c));
    }
}
__global__ void k(int n) {
    #pragma gtap entry
    nq(n, 0, 0u);
}
"""


def test_block_mode_from_the_include():
    out = lower.lower(read("bsort.gtap.cu"), "bsort.gtap.cu")
    assert "(block mode)" in out and "::gtap::entry_block<" in out
    assert "__shared__ unsigned tile[256];" in out  # workgroup memory keeps its declaration
    save = out[out.index("void bsort__save"):out.index("void bsort__load")]
    assert "if (threadIdx.x == 0)" in save and "complete_stores" in save
    assert "struct Shared" in out and "kQueues" not in out[out.index("struct Program"):out.index("struct Program") + 300]


def test_utf8_comments_keep_offsets():
    src = "// 乱数の説明 — コメント\n" + read("fib.gtap.cu").replace("if (n < 2)", "// 葉\n    if (n < 2)")
    out = lower.lower(src, "fib_utf8.cu")
    assert "__gtap_spawn_task(lane, F_fib" in out and "葉" in out


BAD_MORE = {
    "memcpy in task code": """
#pragma gtap function
__device__ void f(int* a, int* b, int n) {
    memcpy(a, b, n * 4);
}
""",
    "inline asm in task code": """
#pragma gtap function
__device__ void f(int* a) {
    asm volatile("s_nop 0");
}
""",
    "result without a later taskwait": """
#pragma gtap function
__device__ int f(int n) {
    if (n < 2) return n;
    int a;
    #pragma gtap task
    a = f(n - 1);
    return a;
}
""",
}


@pytest.mark.parametrize("case", sorted(BAD_MORE))
def test_rejected_unsafe_forms(case):
    with pytest.raises(lower.LoweringError):
        lower.lower(BAD_MORE[case], "bad.cu")
