"""The reference's on-disk formats (graph CSR, key vector, SpMV CSR): the library's readers
and writers against bytes packed independently with struct, against the key files the
reference's own gen_vector wrote (tests/golden), and on malformed files. CPU only."""
import ctypes
import os
import struct

import numpy as np
import pytest

import gtap_amd
import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

pytestmark = pytest.mark.skipif(not os.path.exists(gtap_amd.LIB_PATH), reason="libgtap_amd.so not built")


def test_csr_bytes_and_round_trip(tmp_path):
    # asynchronous_bfs_thread.cu:69-84: int32 N, int32 M, int32 row[N+1], int32 col[M]
    row = np.array([0, 2, 3, 3], dtype=np.int32)
    col = np.array([1, 2, 0], dtype=np.int32)
    p = str(tmp_path / "g.bin")
    gtap_amd.write_csr(p, row, col)
    assert open(p, "rb").read() == struct.pack("<2i", 3, 3) + struct.pack("<4i", *row) + struct.pack("<3i", *col)
    r2, c2 = gtap_amd.read_csr(p)
    assert np.array_equal(r2, row) and np.array_equal(c2, col)
    # a graph file written by hand is read back the same
    q = str(tmp_path / "h.bin")
    with open(q, "wb") as f:
        f.write(struct.pack("<2i", 2, 1) + struct.pack("<3i", 0, 1, 1) + struct.pack("<1i", 1))
    r3, c3 = gtap_amd.read_csr(q)
    assert r3.tolist() == [0, 1, 1] and c3.tolist() == [1]
    # R-MAT graph round trip
    rr, cc = gtap_amd.rmat_csr(10, 4, 3)
    z = str(tmp_path / "rmat.bin")
    gtap_amd.write_csr(z, rr, cc)
    r4, c4 = gtap_amd.read_csr(z)
    assert np.array_equal(r4, rr) and np.array_equal(c4, cc)


@pytest.mark.parametrize("n", [1000, 4097])
def test_key_files_written_by_reference_gen_vector(n):
    # files written by the reference's own gen_vector binary (gen_vector.c:21-36)
    keys = gtap_amd.read_keys(os.path.join(GOLDEN, f"gen_vector_{n}.bin"))
    assert np.array_equal(keys, oracle.gen_keys_u32(n))


def test_key_file_round_trip(tmp_path):
    keys = oracle.gen_keys_u32(777, 5)
    p = str(tmp_path / "k.bin")
    gtap_amd.write_keys(p, keys)
    raw = open(p, "rb").read()
    assert raw[:8] == struct.pack("<Q", 777) and len(raw) == 8 + 4 * 777
    assert np.array_equal(gtap_amd.read_keys(p), keys)


def test_spmv_file_bytes_and_round_trip(tmp_path):
    # spmv_thread.cu:70-86: int32 rows, cols, nnz, row_ptr[rows+1], col_idx[nnz], float64 val[nnz]
    rp, ci, va = gtap_amd.spmv_random(50, 4, 9)
    p = str(tmp_path / "m.bin")
    gtap_amd.write_spmv(p, 50, rp, ci, va)
    raw = open(p, "rb").read()
    nnz = int(rp[-1])
    assert raw[:12] == struct.pack("<3i", 50, 50, nnz)
    assert len(raw) == 12 + 4 * 51 + 12 * nnz
    assert raw[12 + 4 * 51 + 4 * nnz:][:8] == struct.pack("<d", va[0])
    rows, cols, rp2, ci2, va2 = gtap_amd.read_spmv(p)
    assert (rows, cols) == (50, 50)
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci) and np.array_equal(va2, va)


def test_malformed_files_are_rejected(tmp_path):
    bad = str(tmp_path / "bad.bin")
    with open(bad, "wb") as f:
        f.write(struct.pack("<2i", 10, 100))  # header only: truncated
    with pytest.raises(RuntimeError):
        gtap_amd.read_csr(bad)
    with open(bad, "wb") as f:
        f.write(struct.pack("<2i", -1, 0))
    with pytest.raises(RuntimeError):
        gtap_amd.read_csr(bad)
    with open(bad, "wb") as f:
        f.write(struct.pack("<Q", 10) + bytes(8))
    with pytest.raises(RuntimeError):
        gtap_amd.read_keys(bad)
    with open(bad, "wb") as f:
        f.write(struct.pack("<3i", 2, 2, 5))
    with pytest.raises(RuntimeError):
        gtap_amd.read_spmv(bad)
    with pytest.raises(RuntimeError):
        gtap_amd.read_csr(str(tmp_path / "missing.bin"))


def test_spmv_random_matrix_shape():
    # the shape of spmv_thread.cu:89-107: 1 .. 2*avg-1 column draws per row, sorted, unique,
    # values in [-1, 1); deterministic in the seed
    rows, avg = 3000, 10
    rp, ci, va = gtap_amd.spmv_random(rows, avg, 7)
    counts = np.diff(rp)
    assert rp[0] == 0 and counts.min() >= 1 and counts.max() <= 2 * avg - 1
    assert abs(counts.mean() - avg) < 0.5  # draws average avg; duplicates among them are rare here
    for i in range(0, rows, 97):
        c = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(c) > 0) and c.min() >= 0 and c.max() < rows
    assert va.min() >= -1.0 and va.max() < 1.0
    rp2, ci2, va2 = gtap_amd.spmv_random(rows, avg, 7)
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2) and np.array_equal(va, va2)
    rp3, _, _ = gtap_amd.spmv_random(rows, avg, 8)
    assert not np.array_equal(rp, rp3)


def test_tree_input_range_and_determinism():
    a = gtap_amd.tree_input(4096, 1)
    assert a.min() >= -1.0 and a.max() < 1.0 and abs(a.mean()) < 0.05
    assert np.array_equal(a, gtap_amd.tree_input(4096, 1)) and not np.array_equal(a, gtap_amd.tree_input(4096, 2))


@pytest.mark.parametrize("row", [
    [1, 2, 3],        # row[0] != 0
    [0, 3, 2],        # decreasing offsets
    [0, 1, 5],        # last offset beyond the entry count
    [0, 1, 2, 1],     # decreasing at the end
])
def test_malformed_csr_offsets_are_rejected_before_any_launch(tmp_path, row):
    """Row offsets are checked on the host (row[0] == 0, non-decreasing, row[n] == entries)
    before anything is copied or launched, so no device read can run past the columns."""
    row = np.array(row, dtype=np.int32)
    nv = row.size - 1
    col = np.zeros(3, dtype=np.int32)
    depth = np.empty(nv, dtype=np.int32)
    st = gtap_amd.lib().gtap_bfs_host(row.ctypes.data, col.ctypes.data, nv, col.size, 0, depth.ctypes.data, 1, 64,
                                      None)
    assert st != 0
    code = ctypes.c_int(0)
    gtap_amd.lib().gtap_get_runtime_error_code(ctypes.byref(code))
    assert code.value == 11  # GTAP_ERROR_INVALID_ARGUMENT
    x = np.ones(nv, dtype=np.float64)
    y = np.empty(nv, dtype=np.float64)
    val = np.ones(3, dtype=np.float64)
    assert gtap_amd.lib().gtap_spmv_host(row.ctypes.data, col.ctypes.data, val.ctypes.data, x.ctypes.data,
                                         y.ctypes.data, nv, nv, 16, None) != 0
    # the same offsets in a graph file are refused by the reader
    p = str(tmp_path / "bad_rows.bin")
    with open(p, "wb") as f:
        f.write(struct.pack("<2i", nv, 3) + row.tobytes() + col.tobytes())
    with pytest.raises(RuntimeError):
        gtap_amd.read_csr(p)


def test_csr_file_column_out_of_range_is_rejected(tmp_path):
    p = str(tmp_path / "bad_col.bin")
    with open(p, "wb") as f:
        f.write(struct.pack("<2i", 2, 2) + struct.pack("<3i", 0, 1, 2) + struct.pack("<2i", 1, 7))
    with pytest.raises(RuntimeError):
        gtap_amd.read_csr(p)
