"""The reference's own `#pragma gtap` programs, unmodified, lowered by gtap_amd.lower and built
for gfx950 by tests/lowered_ref/build_lowered_ref.py (built here, where the reference tree is;
skipped where a library was not built). Each runs the way its own main() runs it: device
variables bound with hipMemcpyToSymbol, the entry kernel launched by the host, results read
back (examples/fib/fib.cu:26-57, examples/nq/nq.cu:58-88, examples/mergesort/mergesort.cu:99-129,
examples/cilksort/cilksort.cu:215-267, examples/asynchronous_bfs/asynchronous_bfs_thread.cu:86-145,
examples/spmv/spmv_thread.cu:108-212, examples/binary_tree/binary_tree_thread.cu:122-200,
evaluation/3-epaq/fib/fib_queue_{1,3}.cu). Results are checked against the oracle."""
import ctypes
import os

import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lowered_ref", "_lib")
NAMES = ("ref_fib", "ref_nq", "ref_mergesort", "ref_cilksort", "ref_bfs_thread", "ref_bfs_block",
         "ref_spmv_thread", "ref_tree_thread", "ref_fib_queue_1", "ref_fib_queue_3")
INF = oracle.INF_DEPTH
U64 = ctypes.c_uint64


@pytest.fixture(scope="module")
def mods():
    m = {}
    for n in NAMES:
        p = os.path.join(LIBDIR, f"lib{n}.so")
        if os.path.exists(p):
            m[n] = gtap_amd.EntryModule(p, n)
    if not m:
        pytest.skip("no lowered reference program was built (the reference tree was absent at build time)")
    stride = (max(x.record_bytes for x in m.values()) + 63) // 64 * 64
    gtap_amd.gtap_set_config(timeout_ms=20000, record_stride=max(stride, 192))
    assert gtap_amd.gtap_initialize() == 0
    yield m
    assert gtap_amd.gtap_finalize() == 0
    gtap_amd.gtap_set_config()


def need(mods, name):
    if name not in mods:
        pytest.skip(f"lib{name}.so was not built")
    return mods[name]


def run_ok(m, kernel, argtypes, *args):
    assert m.launch(kernel, argtypes, *args) == 0
    code = gtap_amd.gtap_get_runtime_error_code()
    assert code == 0, gtap_amd.ERROR_NAMES.get(code)


@pytest.mark.parametrize("n", [0, 1, 2, 15, 25])
def test_reference_fib(mods, n):
    print(f"case reference fib.cu n={n}", flush=True)
    m = need(mods, "ref_fib")
    m.set_symbol("d_result", -1)
    run_ok(m, "exec_kernel", [ctypes.c_int], n)
    assert m.get_symbol("d_result") == oracle.fib(n)


@pytest.mark.parametrize("n,cutoff", [(1, 7), (4, 7), (6, 2), (8, 7), (10, 3), (11, 7), (12, 5)])
def test_reference_nq_detached_tasks(mods, n, cutoff):
    print(f"case reference nq.cu n={n} cutoff={cutoff}", flush=True)
    m = need(mods, "ref_nq")
    m.set_symbol("d_answer", 0)
    m.set_symbol("d_grid_size", n)
    m.set_symbol("d_cutoff_depth", cutoff)
    run_ok(m, "my_kernel", [])
    # nq.cu counts a placement only in serial_search, which runs below the cutoff row
    # (nq.cu:9-13, 29-32): with n <= cutoff every task stops at row n with no free column,
    # and the program's own answer is 0
    assert m.get_symbol("d_answer") == (oracle.nq(n)[0] if n > cutoff else 0)


def _sort_i32(m, kernel, keys):
    n = keys.size
    a = gtap_amd.DeviceBuffer.from_array(keys)
    t = gtap_amd.DeviceBuffer(max(1, n) * 4)
    try:
        m.set_symbol("g_data", a.ptr, U64)
        m.set_symbol("g_buf", t.ptr, U64)
        run_ok(m, kernel, [ctypes.c_int], n)
        return a.to_array(np.empty(n, dtype=np.int32))
    finally:
        a.free()
        t.free()


@pytest.mark.parametrize("n", [1, 2, 100, 129, 257, 5000, 100_000])
def test_reference_mergesort(mods, n):
    print(f"case reference mergesort.cu n={n}", flush=True)
    keys = oracle.gen_vector(n, 77 + n).view(np.int32)  # the reference's int keys
    assert np.array_equal(_sort_i32(need(mods, "ref_mergesort"), "my_kernel", keys), np.sort(keys))


@pytest.mark.parametrize("n", [1, 63, 64, 65, 256, 1000, 100_001])
def test_reference_cilksort(mods, n):
    print(f"case reference cilksort.cu n={n}", flush=True)
    keys = oracle.gen_vector(n, 91 + n).view(np.int32)
    assert np.array_equal(_sort_i32(need(mods, "ref_cilksort"), "my_kernel", keys), np.sort(keys))


def _graph(kind, scale):
    if kind == "chain":  # the examples' default graph (asynchronous_bfs_thread.cu:57-66)
        n = 1 << scale
        row = np.array([i if i < n else n - 1 for i in range(n + 1)], dtype=np.int32)
        col = np.arange(1, n, dtype=np.int32)
        return row, col
    return gtap_amd.rmat_csr(scale, 16, 3 + scale)


@pytest.mark.parametrize("name,kernel", [("ref_bfs_thread", "my_kernel"), ("ref_bfs_block", "exec_kernel")])
@pytest.mark.parametrize("kind,scale", [("chain", 10), ("rmat", 10), ("rmat", 15)])
def test_reference_bfs(mods, name, kernel, kind, scale):
    print(f"case reference {name} {kind} scale={scale}", flush=True)
    m = need(mods, name)
    row, col = _graph(kind, scale)
    nv = row.size - 1
    dr, dc = gtap_amd.DeviceBuffer.from_array(row), gtap_amd.DeviceBuffer.from_array(col)
    dd = gtap_amd.DeviceBuffer.from_array(np.full(nv, INF, dtype=np.int32))  # memset 0x3f
    try:
        m.set_symbol("g_row_offsets", dr.ptr, U64)
        m.set_symbol("g_col_indices", dc.ptr, U64)
        m.set_symbol("g_depth", dd.ptr, U64)
        m.set_symbol("g_num_vertices", nv)
        run_ok(m, kernel, [ctypes.c_int], 0)
        depth = dd.to_array(np.empty(nv, dtype=np.int32))
    finally:
        for b in (dr, dc, dd):
            b.free()
    assert np.array_equal(depth, oracle.bfs(row, col, 0))


@pytest.mark.parametrize("rows", [1, 17, 1000, 20_000])
def test_reference_spmv(mods, rows):
    print(f"case reference spmv_thread.cu rows={rows}", flush=True)
    m = need(mods, "ref_spmv_thread")
    rp, ci, va = gtap_amd.spmv_random(rows, 24, 5 + rows)
    x = np.linspace(-1.0, 1.0, rows)
    bufs = [gtap_amd.DeviceBuffer.from_array(a) for a in (rp, ci, va, x)] + [gtap_amd.DeviceBuffer(rows * 8)]
    try:
        for sym, b in zip(("g_row_ptr", "g_col_idx", "g_val", "g_x", "g_y"), bufs):
            m.set_symbol(sym, b.ptr, U64)
        run_ok(m, "exec_kernel", [ctypes.c_int], rows)
        y = bufs[-1].to_array(np.empty(rows, dtype=np.float64))
    finally:
        for b in bufs:
            b.free()
    # the reference accumulates sum += v * x; the compiler contracts it to the fma chain the
    # oracle computes (equal bits), any other rounding stays within 1e-12 relative
    np.testing.assert_allclose(y, oracle.spmv(rp, ci, va, x), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("height,mem_ops,iters", [(0, 4, 8), (6, 16, 32), (11, 64, 128)])
def test_reference_binary_tree(mods, height, mem_ops, iters):
    print(f"case reference binary_tree_thread.cu height={height}", flush=True)
    m = need(mods, "ref_tree_thread")
    inp = gtap_amd.tree_input(1 << 12, 21)
    nodes = (1 << (height + 1)) - 1
    di, do = gtap_amd.DeviceBuffer.from_array(inp), gtap_amd.DeviceBuffer(nodes * 8)
    try:
        m.set_symbol("g_input", di.ptr, U64)
        m.set_symbol("g_out", do.ptr, U64)
        m.set_symbol("g_input_n", inp.size)
        run_ok(m, "exec_kernel", [ctypes.c_int, ctypes.c_int, ctypes.c_int], height, mem_ops, iters)
        out = do.to_array(np.empty(nodes, dtype=np.float64))
    finally:
        di.free()
        do.free()
    assert np.array_equal(out, oracle.tree(height, mem_ops, iters, inp))


@pytest.mark.parametrize("name", ["ref_fib_queue_1", "ref_fib_queue_3"])
@pytest.mark.parametrize("n,cutoff", [(2, 2), (20, 2), (30, 13), (33, 13)])
def test_reference_epaq_fib(mods, name, n, cutoff):
    print(f"case reference {name} n={n} cutoff={cutoff}", flush=True)
    m = need(mods, name)
    m.set_symbol("d_cutoff", cutoff)
    m.set_symbol("d_result", -1)
    run_ok(m, "exec_kernel", [ctypes.c_int], n)
    assert m.get_symbol("d_result") == oracle.fib(n)
