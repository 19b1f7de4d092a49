"""gtap_cilksort_noalloc (the merge buffer is the library's static device workspace, no
allocation) against gtap_cilksort with a caller's tmp buffer, through the C ABI: bit-exact
on two calls in a row, the second smaller, and against an independent sort (numpy). Sizes
beyond the workspace are refused before any launch."""
import ctypes

import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 22  # the workspace limit include/gtap.h states


@pytest.fixture(scope="module", autouse=True)
def runtime():
    gtap_amd.gtap_set_config(timeout_ms=20000)
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0


def _sort_with_tmp(keys):
    L = gtap_amd.lib()
    dk, dt = gtap_amd.DeviceBuffer.from_array(keys), gtap_amd.DeviceBuffer(max(4, keys.nbytes))
    try:
        s = gtap_amd.RunStats()
        gtap_amd._check(L.gtap_cilksort(dk.ptr, dt.ptr, keys.size, 0, 0, ctypes.byref(s)), "gtap_cilksort")
        return dk.to_array(np.empty_like(keys))
    finally:
        dk.free()
        dt.free()


def _sort_noalloc(keys):
    L = gtap_amd.lib()
    dk = gtap_amd.DeviceBuffer.from_array(keys)
    try:
        s = gtap_amd.RunStats()
        gtap_amd._check(L.gtap_cilksort_noalloc(dk.ptr, keys.size, 0, 0, ctypes.byref(s)), "gtap_cilksort_noalloc")
        assert s.error_code == 0
        return dk.to_array(np.empty_like(keys))
    finally:
        dk.free()


def test_noalloc_matches_cilksort_on_two_calls():
    # the first call fills the workspace to its limit; the second, smaller one must not be
    # affected by what the first left there
    for n, seed in ((MAX_KEYS, 11), (1000003, 12), (4097, 13)):
        print(f"case cilksort_noalloc n={n} seed={seed}", flush=True)
        keys = oracle.gen_keys_u32(n, seed=seed)
        a = _sort_with_tmp(keys)
        b = _sort_noalloc(keys)
        assert np.array_equal(a, b)
        assert np.array_equal(b, np.sort(keys))


def test_noalloc_refuses_sizes_beyond_the_workspace():
    print("case cilksort_noalloc n=MAX+1 (refused)", flush=True)
    dk = gtap_amd.DeviceBuffer(16)
    try:
        st = gtap_amd.lib().gtap_cilksort_noalloc(dk.ptr, MAX_KEYS + 1, 0, 0, None)
        assert st != 0
        assert gtap_amd.gtap_get_runtime_error_code() == 11  # GTAP_ERROR_INVALID_ARGUMENT
    finally:
        dk.free()
