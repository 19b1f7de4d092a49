"""Programs lowered by gtap_amd.lower, driven as the reference's host programs drive theirs
(examples/fib/fib.cu:26-57): device variables set and read with hipMemcpyToSymbol /
hipMemcpyFromSymbol, the entry kernel launched with <<<grid, block>>> at the co-resident grid
the module reports, hipDeviceSynchronize, then the lhs of `#pragma gtap entry` read back.
Each feature of the source dialect has a program here: a host-set __constant__ cutoff and
queue(expr) with taskwait queue(2) (fib_epaq), plain pointer accesses made coherent and
load_L2 (csort), block mode with workgroup memory (bsort), a statement before the entry and
detached spawns (bfs_l2). Results are checked bit-exactly against the oracle."""
import ctypes

import numpy as np
import pytest

import gtap_amd
import oracle
from gtap_amd import build as gbuild

pytestmark = pytest.mark.gpu

NAMES = ("fib", "nqueens", "msort", "fib_epaq", "fib_cut", "csort", "bsort", "bfs_l2")
INF = oracle.INF_DEPTH


@pytest.fixture(scope="module")
def mods():
    m = {n: gtap_amd.EntryModule(gbuild.lowered_lib(n), n) for n in NAMES}
    stride = (max(x.record_bytes for x in m.values()) + 63) // 64 * 64
    gtap_amd.gtap_set_config(timeout_ms=20000, record_stride=max(stride, 192))
    assert gtap_amd.gtap_initialize() == 0
    yield m
    assert gtap_amd.gtap_finalize() == 0
    gtap_amd.gtap_set_config()


@pytest.mark.parametrize("n", [0, 1, 2, 10, 20, 25])
def test_host_launch_fib(mods, n):
    print(f"case host-launched fib({n})", flush=True)
    m = mods["fib"]
    m.set_symbol("d_fib_result", -1, ctypes.c_longlong)
    assert m.launch("fib_kernel", [ctypes.c_int], n) == 0
    assert gtap_amd.gtap_get_runtime_error_code() == 0
    assert m.get_symbol("d_fib_result", ctypes.c_longlong) == oracle.fib(n)


@pytest.mark.parametrize("n", [4, 8, 10])
def test_host_launch_nqueens(mods, n):
    print(f"case host-launched nqueens({n})", flush=True)
    m = mods["nqueens"]
    assert m.launch("nqueens_kernel", [ctypes.c_int], n) == 0
    assert m.get_symbol("d_nq_result", ctypes.c_longlong) == oracle.nq(n)[0]


@pytest.mark.parametrize("name", ["fib_epaq", "fib_cut"])
@pytest.mark.parametrize("n,cutoff", [(0, 2), (1, 2), (2, 2), (20, 2), (27, 10), (32, 13)])
def test_host_set_cutoff_and_queues(mods, name, n, cutoff):
    print(f"case {name} n={n} cutoff={cutoff}", flush=True)
    m = mods[name]
    m.set_symbol("d_cutoff", cutoff)
    assert m.get_symbol("d_cutoff") == cutoff
    m.set_symbol("d_result", -1, ctypes.c_longlong)
    assert m.launch("fib_kernel", [ctypes.c_int], n) == 0
    assert gtap_amd.gtap_get_runtime_error_code() == 0
    assert m.get_symbol("d_result", ctypes.c_longlong) == oracle.fib(n)


def _sort(m, kernel, keys):
    n = keys.size
    a = gtap_amd.DeviceBuffer.from_array(keys)
    t = gtap_amd.DeviceBuffer(max(1, n) * 4)
    try:
        m.set_symbol("g_keys", a.ptr, ctypes.c_uint64)
        m.set_symbol("g_tmp", t.ptr, ctypes.c_uint64)
        assert m.launch(kernel, [ctypes.c_int], n) == 0
        assert gtap_amd.gtap_get_runtime_error_code() == 0
        return a.to_array(np.empty(n, dtype=np.uint32))
    finally:
        a.free()
        t.free()


@pytest.mark.parametrize("n", [1, 2, 31, 33, 129, 1000, 4097, 65537, 300_001])
def test_coherent_plain_accesses_thread_sort(mods, n):
    print(f"case csort n={n}", flush=True)
    keys = oracle.gen_keys_u32(n, 500 + n)
    assert np.array_equal(_sort(mods["csort"], "csort_kernel", keys), np.sort(keys))


@pytest.mark.parametrize("n", [1, 5, 256, 257, 1000, 4096, 65539, 200_000])
def test_block_mode_sort(mods, n):
    print(f"case bsort n={n}", flush=True)
    keys = oracle.gen_keys_u32(n, 900 + n)
    keys[::7] = keys[0]  # ties across the two halves of every merge
    assert np.array_equal(_sort(mods["bsort"], "bsort_kernel", keys), np.sort(keys))


def _graph(kind, scale):
    if kind == "chain":
        n = 1 << scale
        row = np.append(np.arange(n - 1, dtype=np.int32), [n - 1, n - 1]).astype(np.int32)
        col = np.arange(1, n, dtype=np.int32)
        return row, col
    return gtap_amd.rmat_csr(scale, 16, 7 + scale)


@pytest.mark.parametrize("kind,scale", [("chain", 8), ("rmat", 10), ("rmat", 14), ("rmat", 16)])
def test_bfs_with_pre_entry_statement(mods, kind, scale):
    print(f"case bfs_l2 {kind} scale={scale}", flush=True)
    row, col = _graph(kind, scale)
    nv = row.size - 1
    m = mods["bfs_l2"]
    dr, dc = gtap_amd.DeviceBuffer.from_array(row), gtap_amd.DeviceBuffer.from_array(col)
    dd = gtap_amd.DeviceBuffer.from_array(np.full(nv, INF, dtype=np.int32))
    try:
        m.set_symbol("g_row", dr.ptr, ctypes.c_uint64)
        m.set_symbol("g_col", dc.ptr, ctypes.c_uint64)
        m.set_symbol("g_depth", dd.ptr, ctypes.c_uint64)
        assert m.launch("bfs_kernel", [ctypes.c_int], 0) == 0
        assert gtap_amd.gtap_get_runtime_error_code() == 0
        depth = dd.to_array(np.empty(nv, dtype=np.int32))
    finally:
        for b in (dr, dc, dd):
            b.free()
    assert np.array_equal(depth, oracle.bfs(row, col, 0))


def test_second_launch_needs_a_reset(mods):
    """One launch per gtap_initialize / gtap_reset: a second launch without a reset does
    nothing and reports GTAP_ERROR_NOT_INITIALIZED; after a reset it runs."""
    print("case launch without reset (refused)", flush=True)
    m = mods["fib"]
    assert m.launch("fib_kernel", [ctypes.c_int], 10) == 0
    assert m.get_symbol("d_fib_result", ctypes.c_longlong) == 55
    m.set_symbol("d_fib_result", -7, ctypes.c_longlong)
    assert m.launch("fib_kernel", [ctypes.c_int], 12, reset=False) == 0
    assert gtap_amd.gtap_get_runtime_error_code() == 12
    assert m.get_symbol("d_fib_result", ctypes.c_longlong) == -7
    assert m.launch("fib_kernel", [ctypes.c_int], 12) == 0
    assert gtap_amd.gtap_get_runtime_error_code() == 0
    assert m.get_symbol("d_fib_result", ctypes.c_longlong) == 144


def test_launch_with_a_smaller_grid(mods):
    """Any grid of 256-thread workgroups runs the program (fewer workers than co-resident)."""
    print("case fib on 3 workgroups", flush=True)
    m = mods["fib"]
    for grid in (1, 3, 17):
        assert m.launch("fib_kernel", [ctypes.c_int], 18, grid=grid) == 0
        assert m.get_symbol("d_fib_result", ctypes.c_longlong) == oracle.fib(18)
