/*
 * gtap_oracle.h -- CPU restatement of the GTaP hot-path programs.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (gtap_amd/, include/) may
 * include, link or call this code. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker.
 *
 * Each function cites the reference file:line whose algorithm it restates.
 * How each function is pinned (golden vectors, closed forms, independent
 * library routines) is listed next to it; the tests in tests/test_oracle.py
 * carry out those checks.
 */
#ifndef GTAP_ORACLE_H
#define GTAP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* fib: evaluation/2-comparison/fib/seq_fib.c:5-10 (int64 result instead of int).
 * Pinned: closed form F(n) (Binet / iteration in the tests), reference binary seq_fib. */
int64_t gtap_oracle_fib(int n);
/* Number of task invocations of the fork-join fib tree (root included), counted
 * by walking the tree of evaluation/2-comparison/fib/gtap_fib.cu:10-22.
 * Pinned: closed form 2F(n+1)-1. */
int64_t gtap_oracle_fib_tasks(int n);

/* n-queens: evaluation/2-comparison/nq/seq_nq.c:12-33 bitmask recursion.
 * *solutions = number of placements; *tasks = nodes of the search tree
 * (root + one per legal placement), i.e. the tasks of the spawn-per-legal-column
 * task program with no cutoff.
 * Pinned: OEIS A000170 counts, reference binary seq_nq, brute force over
 * column permutations for small n. */
void gtap_oracle_nq(int n, int64_t* solutions, int64_t* tasks);
/* Search-tree nodes with at least one legal column (the tasks that suspend at the
 * taskwait and resume once). Pinned: brute force in the tests. */
int64_t gtap_oracle_nq_internal(int n);

/* OpenMP-task versions timed as CPU baselines (evaluation/2-comparison/fib/omp_fib.c:5-15,
 * evaluation/2-comparison/nq/omp_nq.c:23-43 without cutoff, taskwait + int64 sum). */
int64_t gtap_oracle_fib_omp(int n, int cutoff);
int64_t gtap_oracle_nq_omp(int n, int cutoff);
int gtap_oracle_omp_threads(void);

/* Key generator: evaluation/util/gen_vector.c:8-18,157-162 (LCG x = x*1103515245+12345
 * mod 2^64 starting from the seed; value = INT32_MIN + (x mod 2^32)).
 * gtap_oracle_gen_vector writes the int32 values; gtap_oracle_gen_keys_u32 writes
 * the same 32-bit patterns as uint32 keys.
 * Pinned: byte-identical to the file the reference gen_vector binary writes. */
void gtap_oracle_gen_vector(uint64_t seed, size_t n, int32_t* out);
void gtap_oracle_gen_keys_u32(uint64_t seed, size_t n, uint32_t* out);

/* cilksort: evaluation/2-comparison/cilksort/gtap_cilksort.cu:18-167 -- four-way
 * split (len12=n/2, len1=len12/2, ...), cilkmerge with a_split=(a_len+1)/2 and the
 * binary search that starts at low=-1 (gtap_cilksort.cu:40-50).
 * Leaves of n <= leaf_max keys are sorted by insertion sort; merges of fewer than
 * merge_cutoff keys are sequential (gtap_cilksort.cu:18-37).
 * Pinned: numpy.sort (an independent sort) and the reference seq_cilksort binary. */
void gtap_oracle_cilksort_u32(uint32_t* arr, uint32_t* tmp, int64_t n, int64_t leaf_max, int64_t merge_cutoff);
/* Binary search of gtap_cilksort.cu:40-50: largest index i with x[i] <= val, or -1. */
int64_t gtap_oracle_binary_search_u32(const uint32_t* x, int64_t len, uint32_t val);

/* mergesort: evaluation/2-comparison/mergesort/seq_mergesort.cpp:38-62.
 * Pinned: numpy.sort. */
void gtap_oracle_mergesort_u32(uint32_t* a, uint32_t* buf, int64_t n);

/* R-MAT edge generator (no counterpart in the reference, which downloads real graphs).
 * SplitMix64 step: s += 0x9E3779B97F4A7C15, z = s, z = (z^(z>>30))*0xBF58476D1CE4E5B9,
 * z = (z^(z>>27))*0x94D049BB133111EB, out = z^(z>>31).
 * Edge e in [0, nedges): s = seed ^ ((e+1)*0xD1B54A32D192ED03); for each of `scale`
 * levels take r = top 32 bits of the next SplitMix64 output and pick the quadrant
 * against the uint32 thresholds ta < tab < tabc (a, a+b, a+b+c scaled by 2^32):
 * (0,0) if r<ta, (0,1) if r<tab, (1,0) if r<tabc, else (1,1); the bits are appended
 * most-significant first. Self loops and duplicates are kept; edges are directed.
 * Pinned: determinism, quadrant frequencies, degree statistics (tests). */
void gtap_oracle_rmat_edges(int scale, int64_t nedges, uint64_t seed,
                            uint32_t ta, uint32_t tab, uint32_t tabc,
                            int32_t* src, int32_t* dst);
/* CSR by a stable counting sort on src (row offsets int32 [nv+1], columns int32 [ne]),
 * the binary layout read at examples/asynchronous_bfs/asynchronous_bfs_thread.cu:69-84. */
void gtap_oracle_csr_from_edges(int32_t nv, int64_t ne, const int32_t* src, const int32_t* dst,
                                int32_t* row, int32_t* col);
/* BFS distances: examples/asynchronous_bfs/asynchronous_bfs_thread.cu:145-174
 * (queue BFS, unreached = 0x3f3f3f3f).
 * Pinned: scipy.sparse.csgraph shortest paths on unit weights (independent code). */
void gtap_oracle_bfs(int32_t nv, const int32_t* row, const int32_t* col, int32_t source, int32_t* depth);

/* Mixed job stream (no counterpart in the reference). Job i starts a SplitMix64 state
 * s = seed ^ ((2i+1)*0xD1B54A32D192ED03) and takes its first two outputs r0, r1:
 *   kind = (r0>>32) % 10 : 0..3 fib, 4..6 n-queens, 7..9 mergesort
 *   fib   n   = 18 + (r1>>32) % 11
 *   nq    N   = 8 + (r1>>32) % 5
 *   sort  len = floor(2^(10 + 10*u)), u = (r1>>11) * 2^-53, clamped to [2^10, 2^20];
 *         key seed = r1 ^ 0x94D049BB133111EB (keys by gtap_oracle_gen_keys_u32).
 * param receives n / N / len. */
void gtap_oracle_jobs(uint64_t seed, int32_t njobs, int32_t* kind, int32_t* param, uint64_t* key_seed);
/* 64-bit FNV-1a over the bytes of a uint32 array (used to compare sorted outputs). */
uint64_t gtap_oracle_hash_u32(const uint32_t* a, int64_t n);
/* A stream sort job's result: sum over positions i of mix((i << 32) ^ a[i]) mod 2^64, with
 * mix the SplitMix64 finalizer (Steele, Lea, Flood, OOPSLA 2014). Every key is weighed at its
 * position, and the terms add in any order (the device sums them in parallel). Pinned: a
 * numpy restatement, and changes under a transposition or one flipped bit. */
uint64_t gtap_oracle_sort_checksum(const uint32_t* a, int64_t n);
/* The whole job stream on the CPU (sequential, or OpenMP parallel-for over the jobs); job
 * results into results[], returns the loop's wall seconds. */
double gtap_oracle_stream_cpu(uint64_t seed, int32_t njobs, int32_t parallel, int64_t* results);

/* SpMV: examples/spmv/spmv_thread.cu:34-46 (one row = sum over the row's entries in column
 * order), with the product-add written as fma(val, x, sum) so the result is one bit pattern.
 * Pinned: dense numpy products on integer-valued matrices (exact), scipy.sparse within
 * rounding on random ones, identity and all-ones closed forms (tests). */
void gtap_oracle_spmv(int32_t rows, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
                      const double* x, double* y);
/* Binary-tree workload: examples/binary_tree/binary_tree_thread.cu:44-120. out[node] for
 * every node of the tree of the given height; the conversion of the (possibly negative)
 * accumulator to uint32 saturates (negative and NaN -> 0, >= 2^32-1 -> 2^32-1).
 * Pinned: a pure-Python restatement with exact rational fma on small trees, and closed forms
 * for mem_ops = compute_iters = 0 (tests). */
void gtap_oracle_tree(int32_t height, int32_t mem_ops, int32_t compute_iters, const double* input, int64_t input_n,
                      double* out);
/* One node's result of the above. */
double gtap_oracle_tree_node(int32_t node, int32_t mem_ops, int32_t compute_iters, const double* input,
                             int64_t input_n);

#ifdef __cplusplus
}
#endif
#endif
