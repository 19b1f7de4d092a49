"""Pins of the CPU oracle (oracle/gtap_oracle.c) -- CPU only.

Each restated function is checked against something that is not its own formula:
the reference's own programs' outputs (tests/golden, written by make_golden.py from the
binaries built out of the reference sources), closed forms, published sequences,
independent library routines (numpy.sort / searchsorted, scipy BFS) and brute force on
tiny inputs. A mistake such as a dropped term, an off-by-one in the split, or the
low=0 binary search of omp_cilksort.cpp:56-66 fails at least one of them.
"""
import itertools
import json
import os
import struct

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
with open(os.path.join(GOLDEN, "reference_outputs.json")) as f:
    REF = json.load(f)

# OEIS A000170 (number of n-queens placements), n = 1..14
A000170 = [1, 0, 0, 2, 10, 4, 40, 92, 352, 724, 2680, 14200, 73712, 365596]


def fib_iter(n):
    a, b = 0, 1
    for _ in range(n):
        a, b = b, a + b
    return a


# ---------------------------------------------------------------- fib

@pytest.mark.parametrize("n", range(0, 33))
def test_fib_matches_reference_program(n):
    assert oracle.fib(n) == REF["seq_fib"][str(n)]


@pytest.mark.parametrize("n", [0, 1, 2, 10, 20, 30, 35])
def test_fib_closed_form(n):
    assert oracle.fib(n) == fib_iter(n)
    # spawned tasks (root included) of the two-spawn task tree
    assert oracle.fib_tasks(n) == 2 * fib_iter(n + 1) - 1


def test_fib_tasks_of_benchmark_config():
    assert oracle.fib_tasks(30) == 2_692_537


def test_fib_omp_matches_serial():
    assert oracle.fib_omp(25, 0) == fib_iter(25)
    assert oracle.fib_omp(30, 15) == fib_iter(30)


# ---------------------------------------------------------------- n-queens

def nq_brute(n):
    """Independent brute force: explicit row-by-row placement with column and
    diagonal sets; returns (solutions, search-tree nodes including the root,
    nodes with at least one child)."""
    nodes = 0
    sols = 0
    internal = 0

    def rec(row, cols, d1, d2):
        nonlocal nodes, sols, internal
        nodes += 1
        if row == n:
            sols += 1
            return
        kids = 0
        for c in range(n):
            if c in cols or (row - c) in d1 or (row + c) in d2:
                continue
            kids += 1
            rec(row + 1, cols | {c}, d1 | {row - c}, d2 | {row + c})
        internal += kids > 0

    rec(0, frozenset(), frozenset(), frozenset())
    return sols, nodes, internal


def nq_perm_count(n):
    return sum(1 for p in itertools.permutations(range(n))
               if len({p[i] - i for i in range(n)}) == n and len({p[i] + i for i in range(n)}) == n)


@pytest.mark.parametrize("n", range(1, 15))
def test_nq_counts_match_oeis(n):
    assert oracle.nq(n)[0] == A000170[n - 1]


@pytest.mark.parametrize("n", range(1, 14))
def test_nq_counts_match_reference_program(n):
    assert oracle.nq(n)[0] == REF["seq_nq"][str(n)]


@pytest.mark.parametrize("n", range(4, 12))
def test_nq_omp_reference_program_agrees(n):
    # the reference's OpenMP-task program (cutoff 3) gives the same counts
    assert REF["omp_nq"][str(n)] == oracle.nq(n)[0]


@pytest.mark.parametrize("n", range(1, 9))
def test_nq_tasks_brute_force(n):
    s, t = oracle.nq(n)
    bs, bt, bi = nq_brute(n)
    assert (s, t) == (bs, bt)
    assert oracle.nq_internal(n) == bi
    if n <= 8:
        assert s == nq_perm_count(n)


def test_nq_task_counts_of_configs():
    assert oracle.nq(8) == (92, 2057)
    assert oracle.nq(12) == (14200, 856189)


def test_nq_omp_sum():
    assert oracle.nq_omp(10, 4) == 724
    assert oracle.nq_omp(12, 2) == 14200


# ---------------------------------------------------------------- keys

@pytest.mark.parametrize("n", [1000, 4097])
def test_gen_vector_byte_identical_to_reference(n):
    raw = open(os.path.join(GOLDEN, f"gen_vector_{n}.bin"), "rb").read()
    assert int.from_bytes(raw[:8], "little") == n
    ref = np.frombuffer(raw[8:], dtype=np.int32)
    assert np.array_equal(oracle.gen_vector(n), ref)
    # uint32 keys are the same bit patterns
    assert np.array_equal(oracle.gen_keys_u32(n), ref.view(np.uint32))


def test_gen_keys_lcg_closed_form():
    # x_k = a^k x_0 + c (a^k - 1)/(a - 1) mod 2^64; key = low 32 bits ^ 0x80000000
    a, c, m = 1103515245, 12345, 1 << 64
    x0 = 123456789
    keys = oracle.gen_keys_u32(50)
    for k in (1, 2, 17, 50):
        ak = a ** k  # exact integer; reduced mod 2^64 only at the end
        xk = (ak * x0 + c * ((ak - 1) // (a - 1))) % m
        assert keys[k - 1] == ((xk & 0xFFFFFFFF) ^ 0x80000000)


# ---------------------------------------------------------------- sorting

SIZES = [0, 1, 2, 3, 63, 64, 65, 255, 256, 257, 4095, 4096, 4097, 10000, 65537]


def _inputs(n, rng):
    yield rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    yield rng.integers(0, 7, n).astype(np.uint32)                     # heavy duplicates
    yield np.arange(n, dtype=np.uint32)                                # sorted
    yield np.arange(n, dtype=np.uint32)[::-1].copy()                   # reversed
    yield np.full(n, 0xFFFFFFFF, dtype=np.uint32)                      # all equal, max key


@pytest.mark.parametrize("n", SIZES)
def test_cilksort_matches_numpy(n):
    rng = np.random.default_rng(n)
    for x in _inputs(n, rng):
        for leaf, mc in ((4096, 256), (64, 256), (1, 3)):
            assert np.array_equal(oracle.cilksort_u32(x, leaf, mc), np.sort(x, kind="stable"))


@pytest.mark.parametrize("n", SIZES)
def test_mergesort_matches_numpy(n):
    rng = np.random.default_rng(1000 + n)
    for x in _inputs(n, rng):
        assert np.array_equal(oracle.mergesort_u32(x), np.sort(x))


def test_binary_search_matches_searchsorted():
    rng = np.random.default_rng(3)
    x = np.sort(rng.integers(0, 100, 257).astype(np.uint32))
    for v in list(range(0, 102)) + [0xFFFFFFFF]:
        assert oracle.binary_search_u32(x, v) == int(np.searchsorted(x, v, side="right")) - 1
    # the search starts at low=-1 (gtap_cilksort.cu:40-50): a pivot below every element gives -1
    assert oracle.binary_search_u32(np.array([5, 6, 7], dtype=np.uint32), 3) == -1
    assert oracle.binary_search_u32(np.array([], dtype=np.uint32), 3) == -1


def test_cilksort_benchmark_keys_sample():
    keys = oracle.gen_keys_u32(1 << 16)
    assert np.array_equal(oracle.cilksort_u32(keys), np.sort(keys))


# ---------------------------------------------------------------- R-MAT + BFS

def test_rmat_deterministic_and_seeded():
    s1, d1 = oracle.rmat_edges(10, 16, seed=5)
    s2, d2 = oracle.rmat_edges(10, 16, seed=5)
    s3, _ = oracle.rmat_edges(10, 16, seed=6)
    assert np.array_equal(s1, s2) and np.array_equal(d1, d2)
    assert not np.array_equal(s1, s3)
    assert s1.min() >= 0 and s1.max() < 1024 and d1.min() >= 0 and d1.max() < 1024


def test_rmat_quadrant_frequencies():
    scale = 14
    src, dst = oracle.rmat_edges(scale, 16, seed=1)
    hi_s = src >> (scale - 1)
    hi_d = dst >> (scale - 1)
    n = src.size
    for (i, j), p in {(0, 0): 0.57, (0, 1): 0.19, (1, 0): 0.19, (1, 1): 0.05}.items():
        f = np.mean((hi_s == i) & (hi_d == j))
        sigma = np.sqrt(p * (1 - p) / n)
        assert abs(f - p) < 5 * sigma, (i, j, f, p)
    # every level uses the same distribution: the low bit too
    lo = np.mean(((src & 1) == 0) & ((dst & 1) == 0))
    assert abs(lo - 0.57) < 5 * np.sqrt(0.57 * 0.43 / n)


def test_csr_is_stable_counting_sort():
    src, dst = oracle.rmat_edges(9, 8, seed=2)
    row, col = oracle.csr_from_edges(512, src, dst)
    assert row[0] == 0 and row[-1] == src.size and np.all(np.diff(row) >= 0)
    order = np.argsort(src, kind="stable")
    assert np.array_equal(col, dst[order])
    assert np.array_equal(np.diff(row), np.bincount(src, minlength=512))


def _scipy_depth(row, col, source):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import shortest_path
    nv = row.size - 1
    g = csr_matrix((np.ones(col.size), col, row), shape=(nv, nv))
    d = shortest_path(g, method="D", directed=True, unweighted=True, indices=source)
    out = np.full(nv, oracle.INF_DEPTH, dtype=np.int64)
    fin = np.isfinite(d)
    out[fin] = d[fin].astype(np.int64)
    return out


@pytest.mark.parametrize("scale,seed,source", [(8, 1, 0), (10, 3, 0), (12, 1, 0), (12, 9, 77)])
def test_bfs_matches_scipy_on_rmat(scale, seed, source):
    row, col = oracle.rmat_csr(scale, 16, seed)
    assert np.array_equal(oracle.bfs(row, col, source), _scipy_depth(row, col, source))


def test_bfs_chain_star_and_unreached():
    n = 50
    row = np.array(list(range(n)) + [n - 1], dtype=np.int32)   # chain 0->1->...->49
    col = np.arange(1, n, dtype=np.int32)
    d = oracle.bfs(row, col, 0)
    assert np.array_equal(d, np.arange(n))
    d = oracle.bfs(row, col, 10)
    assert np.all(d[:10] == oracle.INF_DEPTH) and np.array_equal(d[10:], np.arange(n - 10))
    # star: 0 -> 1..n-1
    row = np.array([0] + [n - 1] * n, dtype=np.int32)
    col = np.arange(1, n, dtype=np.int32)
    d = oracle.bfs(row, col, 0)
    assert d[0] == 0 and np.all(d[1:] == 1)


# ---------------------------------------------------------------- job stream

def test_job_stream_mix():
    kind, param, kseed = oracle.jobs(2024, 4096)
    k2, p2, s2 = oracle.jobs(2024, 4096)
    assert np.array_equal(kind, k2) and np.array_equal(param, p2) and np.array_equal(kseed, s2)
    frac = np.bincount(kind, minlength=3) / 4096
    assert abs(frac[0] - 0.4) < 0.04 and abs(frac[1] - 0.3) < 0.04 and abs(frac[2] - 0.3) < 0.04
    assert param[kind == 0].min() >= 18 and param[kind == 0].max() <= 28
    assert param[kind == 1].min() >= 8 and param[kind == 1].max() <= 12
    lens = param[kind == 2]
    assert lens.min() >= 1024 and lens.max() <= 1 << 20
    # log-uniform: the median of log2(len) is near 15
    assert abs(np.median(np.log2(lens)) - 15.0) < 0.6
    assert np.all(kseed[kind != 2] == 0) and np.all(kseed[kind == 2] != 0)


def test_job_results():
    assert oracle.job_result(0, 20, 0) == 6765
    assert oracle.job_result(1, 8, 0) == 92
    keys = oracle.gen_keys_u32(1500, 99)
    assert oracle.job_result(2, 1500, 99) == oracle.sort_checksum(np.sort(keys))


def _checksum_np(a):
    """The sort checksum restated with numpy's wrapping uint64 arithmetic."""
    a = np.asarray(a, dtype=np.uint64)
    z = (np.arange(a.size, dtype=np.uint64) << np.uint64(32)) ^ a
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    v = int(z.sum(dtype=np.uint64))
    return v - (1 << 64) if v >= 1 << 63 else v


def test_sort_checksum_pins():
    # the finalizer maps 0 to 0: the empty array and a single zero key sum to 0
    assert oracle.sort_checksum(np.array([], dtype=np.uint32)) == 0
    assert oracle.sort_checksum(np.array([0], dtype=np.uint32)) == 0
    for n in (1, 2, 7, 1000, 65537):
        a = np.sort(oracle.gen_keys_u32(n, 3 + n))
        assert oracle.sort_checksum(a) == _checksum_np(a)
    a = np.sort(oracle.gen_keys_u32(4096, 11))
    c = oracle.sort_checksum(a)
    b = a.copy()
    b[[10, 11]] = b[[11, 10]]  # a transposition of two different keys
    assert a[10] != a[11] and oracle.sort_checksum(b) != c
    b = a.copy()
    b[2000] ^= 1 << 17  # one flipped bit
    assert oracle.sort_checksum(b) != c
    # the terms add in any order: the sum over two halves with their positions is the total
    h = a.size // 2
    lo = _checksum_np(a[:h])
    z = (np.arange(h, a.size, dtype=np.uint64) << np.uint64(32)) ^ a[h:].astype(np.uint64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    hi = int((z ^ (z >> np.uint64(31))).sum(dtype=np.uint64))
    assert (lo + hi) % (1 << 64) == c % (1 << 64)


def test_hash_is_fnv1a():
    a = np.array([1, 2, 0xDEADBEEF], dtype=np.uint32)
    h = 0xcbf29ce484222325
    for byte in a.tobytes():
        h ^= byte
        h = (h * 0x100000001b3) % (1 << 64)
    assert oracle.hash_u32(a) == h


# ---------------------------------------------------------------- SpMV and the binary tree

def _dense(rows, cols, row_ptr, col_idx, val):
    A = np.zeros((rows, cols))
    for i in range(rows):
        for k in range(row_ptr[i], row_ptr[i + 1]):
            A[i, col_idx[k]] += val[k]
    return A


def test_oracle_spmv_exact_on_integer_matrices():
    # small integer values: every fma and every dense product is exact, so they agree exactly
    rng = np.random.default_rng(5)
    rows, cols = 40, 30
    counts = rng.integers(0, 6, rows)
    row_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    col_idx = np.concatenate([np.sort(rng.choice(cols, c, replace=False)) for c in counts]).astype(np.int32)
    val = rng.integers(-8, 9, row_ptr[-1]).astype(np.float64)
    x = rng.integers(-8, 9, cols).astype(np.float64)
    y = oracle.spmv(row_ptr, col_idx, val, x)
    assert np.array_equal(y, _dense(rows, cols, row_ptr, col_idx, val) @ x)
    assert np.all(y[counts == 0] == 0.0)


def test_oracle_spmv_random_matrix_against_scipy():
    import scipy.sparse as sp
    import gtap_amd
    if not os.path.exists(gtap_amd.LIB_PATH):
        pytest.skip("libgtap_amd.so not built")
    rp, ci, va = gtap_amd.spmv_random(2000, 12, 3)
    x = np.random.default_rng(1).uniform(-1, 1, 2000)
    y = oracle.spmv(rp, ci, va, x)
    ref = sp.csr_matrix((va, ci, rp), shape=(2000, 2000)) @ x
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)
    # identity matrix: y == x exactly; all-ones x: row sums
    eye = np.arange(2001, dtype=np.int32)
    assert np.array_equal(oracle.spmv(eye, np.arange(2000, dtype=np.int32), np.ones(2000), x), x)


def _py_fma(a: float, b: float, c: float) -> float:
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))  # exact, then one rounding


def _py_tree_node(node, mem_ops, iters, inp):
    def xs(v):
        v &= 0xFFFFFFFF
        v ^= (v << 13) & 0xFFFFFFFF
        v ^= v >> 17
        v ^= (v << 5) & 0xFFFFFFFF
        return v
    seed = xs(node ^ 0x9E3779B9)
    acc = 0.0
    for m in range(mem_ops):
        acc += float(inp[xs(seed + m * 747796405) & (len(inp) - 1)])
    sat = 0 if not acc > 0 else (0xFFFFFFFF if acc >= 4294967295.0 else int(acc))
    mix = ((0x9E3779B9 ^ node) + sat) & 0xFFFFFFFF
    for it in range(iters):
        y = _py_fma(float(xs(seed + it * 747796405)), 1.0000001192092896, 0.9999999403953552)
        mix ^= struct.unpack("<Q", struct.pack("<d", y))[0] & 0xFFFFFFFF
    return float(mix)


def test_oracle_tree_against_exact_python():
    inp = np.random.default_rng(2).uniform(-1, 1, 256)
    out = oracle.tree(4, 7, 9, inp)
    assert out.size == 31
    for node in range(31):
        assert out[node] == _py_tree_node(node, 7, 9, inp)


def test_oracle_tree_closed_forms_and_saturation():
    inp = np.full(16, -0.5)
    # no work: out = (0x9e3779b9 ^ node) as a double
    out = oracle.tree(3, 0, 0, inp)
    assert out.tolist() == [float(0x9E3779B9 ^ v) for v in range(15)]
    # a negative accumulator converts to 0 (saturation), so it adds nothing
    assert oracle.tree_node(5, 10, 0, inp) == float(0x9E3779B9 ^ 5)
    # a large positive one saturates at 2^32 - 1
    big = np.full(16, 1e12)
    assert oracle.tree_node(5, 3, 0, big) == float(((0x9E3779B9 ^ 5) + 0xFFFFFFFF) & 0xFFFFFFFF)


def test_stream_cpu_baseline_matches_job_results():
    """The CPU stream baseline computes every job like oracle.job_result, sequentially and
    with the OpenMP loop over the jobs."""
    t1, seq = oracle.stream_cpu(7, 40, False)
    t2, par = oracle.stream_cpu(7, 40, True)
    kind, param, kseed = oracle.jobs(7, 40)
    exp = [oracle.job_result(int(k), int(p), int(s)) for k, p, s in zip(kind, param, kseed)]
    assert seq.tolist() == exp and par.tolist() == exp
    assert t1 > 0 and t2 > 0
