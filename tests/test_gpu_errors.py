"""Device error paths (GPU): a program that spawns to, or resumes on, an EPAQ queue that does
not exist, and one that returns before the taskwait joining its child, make the runtime
record the reference's error code (runtime/detail/common/gtap_runtime_error.cuh:7-18:
1 invalid queue index, 2 invalid queue index after join, 8 invalid taskwait) and end the
launch normally; the next launch on the same runtime is exact. A task-id pool too small for
the program reports 6 (pool low headroom), and the runtime then runs a program that fits."""
import ctypes

import pytest

import gtap_amd
import oracle
from gtap_amd import build as gbuild

pytestmark = pytest.mark.gpu
I = ctypes.c_int


@pytest.fixture(scope="module")
def mods():
    m = {n: gtap_amd.EntryModule(gbuild.lowered_lib(n), n) for n in ("errors", "fib")}
    stride = (max(x.record_bytes for x in m.values()) + 63) // 64 * 64
    gtap_amd.gtap_set_config(timeout_ms=20000, record_stride=max(stride, 192))
    assert gtap_amd.gtap_initialize() == 0
    yield m
    assert gtap_amd.gtap_finalize() == 0
    gtap_amd.gtap_set_config()


def report():
    r = gtap_amd.ErrorReport()
    assert gtap_amd.lib().gtap_get_runtime_error_report(ctypes.byref(r)) == 0
    return r


def clean_chain(m):
    m.set_symbol("d_out", -1, ctypes.c_longlong)
    assert m.launch("chain_kernel", [I, I, I], 6, 1, 0) == 0
    assert gtap_amd.gtap_get_runtime_error_code() == 0
    assert m.get_symbol("d_out", ctypes.c_longlong) == 7


@pytest.mark.parametrize("q,w,code", [(5, 0, 1), (0, 7, 2)])
def test_invalid_queue_index(mods, q, w, code):
    print(f"case chain with spawn queue {q}, resume queue {w} (error {code} expected)", flush=True)
    m = mods["errors"]
    assert m.launch("chain_kernel", [I, I, I], 6, q, w) == 0  # the kernel ends normally
    assert gtap_amd.gtap_get_runtime_error_code() == code
    r = report()
    assert r.valid == 1 and r.code == code and r.queue_idx == (q if code == 1 else w) and r.limit == 2
    clean_chain(m)


def test_joined_children_beyond_deque_capacity(mods):
    """Joined children stay private until their parent's round ends, so a task that joins more
    children than a deque holds (32768 slots, 31744 usable) ends the launch with a queue
    overflow (3), where the reference traps (gtap_runtime_thread.cuh:42-45); the same program
    within the capacity, and the next launches, are exact."""
    m = mods["errors"]
    print("case wide(20000) (fits)", flush=True)
    m.set_symbol("d_out", -1, ctypes.c_longlong)
    assert m.launch("wide_kernel", [I], 20000) == 0
    assert gtap_amd.gtap_get_runtime_error_code() == 0
    assert m.get_symbol("d_out", ctypes.c_longlong) == 20000
    print("case wide(40000) (queue overflow 3 expected)", flush=True)
    assert m.launch("wide_kernel", [I], 40000) == 0
    assert gtap_amd.gtap_get_runtime_error_code() == 3
    r = report()
    assert r.valid == 1 and r.code == 3 and r.limit == 32768 - 1024
    clean_chain(m)


def test_joined_child_left_behind(mods):
    print("case task returning before its taskwait (error 8 expected)", flush=True)
    m = mods["errors"]
    assert m.launch("leave_kernel", [I], 5) == 0
    assert gtap_amd.gtap_get_runtime_error_code() == 8
    clean_chain(m)
    f = mods["fib"]
    assert f.launch("fib_kernel", [I], 20) == 0
    assert f.get_symbol("d_fib_result", ctypes.c_longlong) == oracle.fib(20)


def test_library_launch_returns_the_error(mods):
    print("case library launch of the invalid queue (status and report)", flush=True)
    L = mods["errors"].L
    f = L.gtap_entry_chain_kernel
    f.argtypes = [I, I, I, ctypes.POINTER(ctypes.c_longlong), ctypes.c_void_p]
    f.restype = ctypes.c_int
    r = ctypes.c_longlong(0)
    s = gtap_amd.RunStats()
    assert f(6, 3, 0, ctypes.byref(r), ctypes.addressof(s)) != 0
    assert s.error_code == 1 and gtap_amd.gtap_get_runtime_error_code() == 1
    assert f(6, 1, 1, ctypes.byref(r), ctypes.addressof(s)) == 0
    assert r.value == 7 and s.error_code == 0


def test_pool_exhaustion_then_a_program_that_fits(mods):
    """Four wavefront workers and the smallest pool (4096 task records): fib(27) holds more
    live tasks than that, the refill finds the pool empty and reports error 6; fib(12) then
    runs exactly on the same runtime."""
    print("case fib(27) on a 4096-record pool (error 6 expected)", flush=True)
    assert gtap_amd.gtap_finalize() == 0
    gtap_amd.gtap_set_config(timeout_ms=20000, pool_bytes=4096 * 192, blocks=1)
    try:
        assert gtap_amd.gtap_initialize() == 0
        with pytest.raises(gtap_amd.GtapError) as e:
            gtap_amd.fib(27)
        assert e.value.code == 6
        print("case fib(12) after the exhaustion", flush=True)
        v, s = gtap_amd.fib(12)
        assert v == oracle.fib(12) and s["error_code"] == 0
    finally:
        assert gtap_amd.gtap_finalize() == 0
        stride = (max(x.record_bytes for x in mods.values()) + 63) // 64 * 64
        gtap_amd.gtap_set_config(timeout_ms=20000, record_stride=max(stride, 192))
        assert gtap_amd.gtap_initialize() == 0
