"""CPU oracle for the GTaP hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker, never the product. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import it.
The product path (``gtap_amd``) never imports, links or calls anything here.

It wraps ``_build/libgtap_oracle.so`` (built from ``gtap_oracle.c`` by ``make``
in this directory) and the reference's own CPU programs built into ``_ref/``
by ``make ref`` (only where the reference sources are present).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgtap_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REFERENCE_ROOT = "/root/reference"

# R-MAT thresholds (a, a+b, a+b+c) = (0.57, 0.76, 0.95) scaled by 2^32, rounded.
RMAT_THRESHOLDS = (2448131359, 3264175145, 4080218931)
INF_DEPTH = 0x3F3F3F3F

_lib: Optional[ctypes.CDLL] = None


def build(force: bool = False, ref: bool = True) -> None:
    """Compile the oracle (and the reference CPU programs when the sources exist)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    if ref and os.path.isdir(REFERENCE_ROOT):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REFERENCE_ROOT}"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = ctypes.CDLL(LIB_PATH)
        i32, i64, u32, u64, sz = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t
        vp = ctypes.c_void_p
        L.gtap_oracle_fib.argtypes = [ctypes.c_int]
        L.gtap_oracle_fib.restype = i64
        L.gtap_oracle_fib_tasks.argtypes = [ctypes.c_int]
        L.gtap_oracle_fib_tasks.restype = i64
        L.gtap_oracle_nq.argtypes = [ctypes.c_int, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.gtap_oracle_nq.restype = None
        L.gtap_oracle_nq_internal.argtypes = [ctypes.c_int]
        L.gtap_oracle_nq_internal.restype = i64
        L.gtap_oracle_fib_omp.argtypes = [ctypes.c_int, ctypes.c_int]
        L.gtap_oracle_fib_omp.restype = i64
        L.gtap_oracle_nq_omp.argtypes = [ctypes.c_int, ctypes.c_int]
        L.gtap_oracle_nq_omp.restype = i64
        L.gtap_oracle_omp_threads.argtypes = []
        L.gtap_oracle_omp_threads.restype = ctypes.c_int
        L.gtap_oracle_gen_vector.argtypes = [u64, sz, vp]
        L.gtap_oracle_gen_keys_u32.argtypes = [u64, sz, vp]
        L.gtap_oracle_cilksort_u32.argtypes = [vp, vp, i64, i64, i64]
        L.gtap_oracle_binary_search_u32.argtypes = [vp, i64, u32]
        L.gtap_oracle_binary_search_u32.restype = i64
        L.gtap_oracle_mergesort_u32.argtypes = [vp, vp, i64]
        L.gtap_oracle_rmat_edges.argtypes = [ctypes.c_int, i64, u64, u32, u32, u32, vp, vp]
        L.gtap_oracle_csr_from_edges.argtypes = [i32, i64, vp, vp, vp, vp]
        L.gtap_oracle_bfs.argtypes = [i32, vp, vp, i32, vp]
        L.gtap_oracle_jobs.argtypes = [u64, i32, vp, vp, vp]
        L.gtap_oracle_hash_u32.argtypes = [vp, i64]
        L.gtap_oracle_hash_u32.restype = u64
        L.gtap_oracle_sort_checksum.argtypes = [vp, i64]
        L.gtap_oracle_sort_checksum.restype = u64
        L.gtap_oracle_stream_cpu.argtypes = [u64, i32, i32, vp]
        L.gtap_oracle_stream_cpu.restype = ctypes.c_double
        L.gtap_oracle_spmv.argtypes = [i32, vp, vp, vp, vp, vp]
        L.gtap_oracle_tree.argtypes = [i32, i32, i32, vp, i64, vp]
        L.gtap_oracle_tree_node.argtypes = [i32, i32, i32, vp, i64]
        L.gtap_oracle_tree_node.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def fib(n: int) -> int:
    return int(lib().gtap_oracle_fib(n))


def fib_tasks(n: int) -> int:
    return int(lib().gtap_oracle_fib_tasks(n))


def nq(n: int) -> tuple[int, int]:
    s, t = ctypes.c_int64(), ctypes.c_int64()
    lib().gtap_oracle_nq(n, ctypes.byref(s), ctypes.byref(t))
    return int(s.value), int(t.value)


def nq_internal(n: int) -> int:
    return int(lib().gtap_oracle_nq_internal(n))


def fib_omp(n: int, cutoff: int = 0) -> int:
    return int(lib().gtap_oracle_fib_omp(n, cutoff))


def nq_omp(n: int, cutoff: int = 64) -> int:
    return int(lib().gtap_oracle_nq_omp(n, cutoff))


def omp_threads() -> int:
    return int(lib().gtap_oracle_omp_threads())


def gen_keys_u32(n: int, seed: int = 123456789) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    lib().gtap_oracle_gen_keys_u32(seed, n, _p(out))
    return out


def gen_vector(n: int, seed: int = 123456789) -> np.ndarray:
    out = np.empty(n, dtype=np.int32)
    lib().gtap_oracle_gen_vector(seed, n, _p(out))
    return out


def cilksort_u32(keys: np.ndarray, leaf_max: int = 4096, merge_cutoff: int = 256) -> np.ndarray:
    a = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    tmp = np.empty_like(a)
    lib().gtap_oracle_cilksort_u32(_p(a), _p(tmp), a.size, leaf_max, merge_cutoff)
    return a


def binary_search_u32(x: np.ndarray, val: int) -> int:
    x = np.ascontiguousarray(x, dtype=np.uint32)
    return int(lib().gtap_oracle_binary_search_u32(_p(x), x.size, val))


def mergesort_u32(keys: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    buf = np.empty_like(a)
    lib().gtap_oracle_mergesort_u32(_p(a), _p(buf), a.size)
    return a


def rmat_edges(scale: int, edge_factor: int = 16, seed: int = 1,
               thresholds: tuple[int, int, int] = RMAT_THRESHOLDS) -> tuple[np.ndarray, np.ndarray]:
    ne = (1 << scale) * edge_factor
    src = np.empty(ne, dtype=np.int32)
    dst = np.empty(ne, dtype=np.int32)
    lib().gtap_oracle_rmat_edges(scale, ne, seed, thresholds[0], thresholds[1], thresholds[2], _p(src), _p(dst))
    return src, dst


def csr_from_edges(nv: int, src: np.ndarray, dst: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    row = np.empty(nv + 1, dtype=np.int32)
    col = np.empty(src.size, dtype=np.int32)
    lib().gtap_oracle_csr_from_edges(nv, src.size, _p(src), _p(dst), _p(row), _p(col))
    return row, col


def rmat_csr(scale: int, edge_factor: int = 16, seed: int = 1) -> tuple[np.ndarray, np.ndarray]:
    src, dst = rmat_edges(scale, edge_factor, seed)
    return csr_from_edges(1 << scale, src, dst)


def bfs(row: np.ndarray, col: np.ndarray, source: int = 0) -> np.ndarray:
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nv = row.size - 1
    depth = np.empty(nv, dtype=np.int32)
    lib().gtap_oracle_bfs(nv, _p(row), _p(col), source, _p(depth))
    return depth


def jobs(seed: int, njobs: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    kind = np.empty(njobs, dtype=np.int32)
    param = np.empty(njobs, dtype=np.int32)
    kseed = np.empty(njobs, dtype=np.uint64)
    lib().gtap_oracle_jobs(seed, njobs, _p(kind), _p(param), _p(kseed))
    return kind, param, kseed


def hash_u32(a: np.ndarray) -> int:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return int(lib().gtap_oracle_hash_u32(_p(a), a.size))


def sort_checksum(a: np.ndarray) -> int:
    """A stream sort job's result for its sorted keys `a` (as an int64 bit pattern)."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    v = int(lib().gtap_oracle_sort_checksum(_p(a) if a.size else None, a.size))
    return v - (1 << 64) if v >= 1 << 63 else v


def stream_cpu(seed: int, njobs: int, parallel: bool) -> tuple[float, np.ndarray]:
    """The whole job stream on this host's CPU (sequential, or OpenMP parallel-for over the
    jobs): (wall seconds, results)."""
    res = np.empty(njobs, dtype=np.int64)
    dt = lib().gtap_oracle_stream_cpu(seed, njobs, 1 if parallel else 0, _p(res))
    return float(dt), res


def job_result(kind: int, param: int, key_seed: int) -> int:
    """Expected result of one stream job: fib value, n-queens count, or the checksum of the
    sorted keys (as the int64 the device returns)."""
    if kind == 0:
        return fib(param)
    if kind == 1:
        return nq(param)[0]
    keys = gen_keys_u32(param, key_seed)
    return sort_checksum(mergesort_u32(keys))


def ref_binary(name: str) -> Optional[str]:
    p = os.path.join(REF_DIR, name)
    return p if os.path.exists(p) else None


def spmv(row_ptr: np.ndarray, col_idx: np.ndarray, val: np.ndarray, x: np.ndarray) -> np.ndarray:
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(row_ptr.size - 1, dtype=np.float64)
    lib().gtap_oracle_spmv(row_ptr.size - 1, _p(row_ptr), _p(col_idx) if col_idx.size else None,
                           _p(val) if val.size else None, _p(x) if x.size else None, _p(y) if y.size else None)
    return y


def tree(height: int, mem_ops: int, compute_iters: int, inp: np.ndarray) -> np.ndarray:
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    out = np.empty((1 << (height + 1)) - 1, dtype=np.float64)
    lib().gtap_oracle_tree(height, mem_ops, compute_iters, _p(inp), inp.size, _p(out))
    return out


def tree_node(node: int, mem_ops: int, compute_iters: int, inp: np.ndarray) -> float:
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    return lib().gtap_oracle_tree_node(node, mem_ops, compute_iters, _p(inp), inp.size)
