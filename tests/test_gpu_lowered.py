"""Programs lowered from `#pragma gtap` sources by gtap_amd.lower, run on the GPU through
their generated entry points, checked bit-exactly against the oracle (fib, n-queens) and a
plain sort (merge sort)."""
import ctypes

import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def runtime():
    need = max(gtap_amd.lowered(n)[f"gtap_lowered_{n}_record_bytes"]() for n in ("fib", "fib32", "nqueens", "msort"))
    stride = (need + 63) // 64 * 64
    gtap_amd.gtap_set_config(timeout_ms=15000, record_stride=stride)
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0


def entry(lib, name, argtypes):
    f = getattr(lib, name)
    f.argtypes = argtypes
    f.restype = ctypes.c_int
    return f


@pytest.mark.parametrize("n", [0, 1, 2, 5, 10, 15, 20, 25])
def test_lowered_fib(n):
    print(f"case lowered fib({n})", flush=True)
    f = entry(gtap_amd.lowered("fib"), "gtap_entry_fib_kernel",
              [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.c_void_p])
    r = ctypes.c_longlong(-1)
    s = gtap_amd.RunStats()
    assert f(n, ctypes.byref(r), ctypes.addressof(s)) == 0
    assert r.value == oracle.fib(n)
    assert s.tasks_executed == (3 * oracle.fib(n + 1) - 2 if n >= 1 else 1)


@pytest.mark.parametrize("n", [0, 1, 2, 5, 10, 20, 27])
def test_lowered_fib32_additive_join(n):
    """The 32-bit fib (the reference's fib.cu shape) is lowered to the additive join."""
    print(f"case lowered fib32({n})", flush=True)
    f = entry(gtap_amd.lowered("fib32"), "gtap_entry_fib32_kernel",
              [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p])
    r = ctypes.c_int(-1)
    s = gtap_amd.RunStats()
    assert f(n, ctypes.byref(r), ctypes.addressof(s)) == 0
    assert r.value == oracle.fib(n)
    assert s.tasks_executed == (3 * oracle.fib(n + 1) - 2 if n >= 1 else 1)


@pytest.mark.parametrize("n", [1, 4, 6, 8, 10, 11])
def test_lowered_nqueens(n):
    print(f"case lowered nqueens({n})", flush=True)
    f = entry(gtap_amd.lowered("nqueens"), "gtap_entry_nqueens_kernel",
              [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.c_void_p])
    r = ctypes.c_longlong(-1)
    s = gtap_amd.RunStats()
    assert f(n, ctypes.byref(r), ctypes.addressof(s)) == 0
    sols, tasks = oracle.nq(n)
    assert r.value == sols
    assert s.tasks_spawned + 1 == tasks
    assert s.tasks_executed == tasks + oracle.nq_internal(n)


@pytest.mark.parametrize("n", [0, 1, 31, 33, 1000, 4097, 100_000])
def test_lowered_msort(n):
    print(f"case lowered msort({n})", flush=True)
    f = entry(gtap_amd.lowered("msort"), "gtap_entry_msort_kernel",
              [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p])
    keys = oracle.gen_keys_u32(n, 17 + n)
    a = gtap_amd.DeviceBuffer.from_array(keys)
    tmp = gtap_amd.DeviceBuffer(max(1, n) * 4)
    try:
        assert f(a.ptr, tmp.ptr, n, None) == 0
        out = a.to_array(np.empty(n, dtype=np.uint32))
        assert np.array_equal(out, np.sort(keys))
    finally:
        a.free()
        tmp.free()
