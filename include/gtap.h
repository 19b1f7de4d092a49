/*
 * gtap.h -- C ABI of the MI355X-native GTaP runtime (libgtap_amd.so).
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and returns an
 * int status that is 0 on success and a hipError_t value otherwise (the reference's
 * functions return cudaError_t). A device-side runtime error does not trap: the
 * persistent kernel records the error, every worker leaves its loop, and the call
 * returns hipErrorLaunchFailure; gtap_get_runtime_error_code() and
 * gtap_print_runtime_error_report() then give the GTaP error code and report.
 *
 * Citations name the reference interface each entry point replaces.
 */
#ifndef GTAP_H
#define GTAP_H

#include <stdint.h>

#include "../gtap_amd/csrc/gtap_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ lifecycle */

/* Replaces gtap_initialize (runtime/detail/thread/gtap_runtime_thread.cuh:338-340,
 * runtime/detail/block/gtap_runtime_block.cuh:155-157): allocates every pool. */
int gtap_initialize(void);
/* Replaces gtap_finalize (gtap_runtime_thread.cuh:342-344). */
int gtap_finalize(void);
/* Replaces gtap_reset (gtap_runtime_thread.cuh:439-441): clears state, keeps allocations, and
 * arms the state for one host-launched entry kernel (as gtap_initialize does). */
int gtap_reset(void);
/* Registers a module written by gtap_amd.lower whose `__global__` entry kernels the host
 * launches itself (the reference's `kernel<<<GTAP_GRID_SIZE, GTAP_BLOCK_SIZE>>>(...)` with
 * `#pragma gtap entry` inside, examples/fib/fib.cu:21-46). gtap_initialize and gtap_reset
 * reset the runtime state, arm it for one such launch, and call `setup` with the runtime
 * descriptor (a module loaded after gtap_initialize gets it at the next gtap_reset); the
 * generated code calls this from a static initializer. No reference
 * counterpart: the reference runtime is compiled into the program's own translation unit. */
int gtap_register_entry_module(gtap_entry_setup_fn setup);
/* Runtime knobs that the reference fixes at compile time with GTAP_* macros
 * (runtime/detail/common/gtap_config.cuh:6-30; examples/README.md:97-105).
 * Must precede gtap_initialize. Zero fields take defaults. */
int gtap_set_config(const gtap_config* cfg);
int gtap_get_config(gtap_config* out);

/* ------------------------------------------------------------ errors */

/* Replaces gtap_synchronize (runtime/detail/common/gtap_runtime_error.cuh:246-249). */
int gtap_synchronize(void);
/* Replaces gtap_check_runtime_error (gtap_runtime_error.cuh:219-234). */
int gtap_check_runtime_error(void);
/* Replaces gtap_get_runtime_error_code (gtap_runtime_error.cuh:38-40). */
int gtap_get_runtime_error_code(int* error_code);
/* Replaces gtap_print_runtime_error_report (gtap_runtime_error.cuh:206-217); returns 1 when
 * a report was printed. */
int gtap_print_runtime_error_report(void);
/* Replaces gtap_get_runtime_error_string (gtap_runtime_error.cuh:187-192). */
const char* gtap_get_runtime_error_string(int error_code);
/* The report itself (the reference keeps it in a static host pointer,
 * gtap_runtime_error.cuh:36). */
int gtap_get_runtime_error_report(gtap_error_report* out);

/* ------------------------------------------------------------ profiling / statistics */

/* Replaces gtap_visualize_profile (runtime/detail/thread/gtap_profile_thread.cuh:355-358):
 * writes the reference's four CSV files (<app>_warp_timeline_having_task.csv,
 * _warp_statistics_having_task.csv, _warp_timeline_working.csv, _warp_statistics_working.csv,
 * same columns) for the last launch when timeline profiling is on, and always
 * <app>_warp_counters.csv with the scheduler counters per worker. */
void gtap_visualize_profile(const char* app_name);
/* The two views on their own, each printing its text timeline of the first ten workers and
 * writing its two CSV files: replace visualize_having_task_time and visualize_working_time
 * (runtime/detail/thread/gtap_profile_thread.cuh:110, 270), which applications call directly
 * (examples/cilksort/cilksort.cu:262). */
void gtap_visualize_having_task_time(const char* app_name);
void gtap_visualize_working_time(const char* app_name);
/* Timeline profiling on/off for the following launches, up to max_samples stamps per worker
 * and kind (0 -> 4096); replaces the PROFILE / MAX_PROFILE_DATA compile-time switches
 * (runtime/detail/common/gtap_config.cuh:22-26). */
int gtap_set_profiling(int enable, uint32_t max_samples);
/* Shader-clock cycles of the last thread-mode launch's round phases, summed over its workers:
 * out[0] acquire, [1] record loads, [2] task bodies, [3] record writes issued, [4] their
 * completion wait, [5] join atomics, [6] continuations and publish, [7] idle tries, and
 * out[8] rounds. Needs timeline profiling on for that launch (the phase stamps are in the
 * profiling build only); n >= 9. No reference counterpart. */
int gtap_get_round_phases(uint64_t* out, uint32_t n);
/* Statistics of the last persistent launch (no reference counterpart; the reference
 * prints "Execution time" around its launch, examples/fib/fib.cu:36-51). */
int gtap_get_last_stats(gtap_run_stats* out);
/* Per-wave counters of the last launch: 16 uint64 per worker, in the order tasks executed,
 * tasks spawned, rounds, pops, steals, steal attempts, jobs admitted, pool refills, records
 * published, records pushed out of the LDS window, suspended taskwaits, the 100 MHz
 * wall-clock ticks at which the worker started and stopped, then 3 unused. */
int gtap_get_worker_stats(uint64_t* out, uint32_t max_workers, uint32_t* nworkers_out);
int gtap_device_info(char* name, int name_len, int* num_cus, uint64_t* hbm_bytes);

/* Measured peaks of this GPU for roofline fractions (no reference counterpart):
 * streaming-copy HBM bandwidth (read + write bytes, GB/s) and agent-scope relaxed atomic
 * fetch-add throughput on distinct 64-byte words (Gop/s). Best of `reps` launches. */
int gtap_bench_hbm_copy(uint64_t bytes, int reps, double* gbps);
int gtap_bench_atomics(int reps, double* gops);
/* The same with every lane adding to one of `words` words (contended when words is small). */
int gtap_bench_atomics_contended(int reps, uint32_t words, double* gops);
/* Random 4-byte accesses over a table of table_bytes: out[0] plain loads, out[1] agent-scope
 * fetch-min, in G accesses/s (the request-rate ceiling of the BFS expansion). No reference
 * counterpart. */
int gtap_bench_random_access(uint64_t table_bytes, int reps, double* out);
/* Time per publication per wave (ns) with `waves` concurrent waves: out_ns[0] an agent-scope
 * release fence, out_ns[1] an agent-scope acquire fence, out_ns[2] a store-completion wait
 * (the order this runtime publishes with). No reference counterpart; a measurement. */
int gtap_bench_fences(uint32_t waves, int iters, double* out_ns);

/* ------------------------------------------------------------ device memory helpers */
/* For callers that only hold host pointers (ctypes, plain C). Each completes before it returns
 * and after the runtime's earlier launches (which run on the runtime's own stream). */
int gtap_device_malloc(void** ptr, uint64_t bytes);
int gtap_device_free(void* ptr);
int gtap_memcpy_htod(void* dst, const void* src, uint64_t bytes);
int gtap_memcpy_dtoh(void* dst, const void* src, uint64_t bytes);
int gtap_memcpy_dtod(void* dst, const void* src, uint64_t bytes);
int gtap_memset_device(void* dst, int value, uint64_t bytes);

/* ------------------------------------------------------------ task programs */
/* Each call is one persistent launch of the program's lowered task functions, with the
 * root task taken from the job table (the `#pragma gtap entry` of the reference). */

/* fib(n) with an int64 result; replaces the program of
 * evaluation/2-comparison/fib/gtap_fib.cu:10-27 + main (examples/fib/fib.cu:26-57). */
int gtap_fib(int n, int64_t* result, gtap_run_stats* stats);

/* Number of n-queens placements, n in [1, 16]: one task per legal placement, each
 * returning its int64 count through a taskwait; replaces the program of
 * evaluation/2-comparison/nq/gtap_nq.cu:25-99 (which counts with a global atomicAdd).
 * The counts are summed by the additive join (each child's count rides on its join atomic). */
int gtap_nqueens(int n, int64_t* result, gtap_run_stats* stats);
/* The same tasks and count with result slots: each child stores its count into the parent's
 * frame, then counts the join down (no reference counterpart; work-stealing scheduler only). */
int gtap_nqueens_slots(int n, int64_t* result, gtap_run_stats* stats);

/* Asynchronous BFS from `source` over a CSR graph in device memory (row: nv+1 int32,
 * col: ne int32; depth: nv int32 out, 0x3f3f3f3f = unreached). num_queues 1 = one queue,
 * 4 = EPAQ by out-degree bucket (<8, <64, <512, >=512), the last bucket's ranges scanned by a
 * whole wavefront each; `split` = most edges per lane range task (0 -> 64; longer ranges are
 * cut into up to 8 pieces). Replaces examples/asynchronous_bfs/asynchronous_bfs_thread.cu:19-141 and
 * evaluation/2-comparison/bfs/gtap_thread_bfs.cu:21-60. */
int gtap_bfs(const int32_t* row, const int32_t* col, int32_t nv, int64_t ne, int32_t source, int32_t* depth,
             int32_t num_queues, int32_t split, gtap_run_stats* stats);
/* The same with host buffers (copied in and out around the launch). */
int gtap_bfs_host(const int32_t* row, const int32_t* col, int32_t nv, int64_t ne, int32_t source,
                  int32_t* depth_out, int32_t num_queues, int32_t split, gtap_run_stats* stats);

/* Cilksort of n uint32 keys in place, block-cooperative (one task per 256-thread
 * workgroup): keys and tmp are device buffers of n keys; leaves of <= leaf keys (0 -> 4096,
 * max 4096) are sorted in LDS by one workgroup; merges of <= merge_leaf keys (0 -> 32768)
 * are merged by one workgroup. Replaces evaluation/2-comparison/cilksort/gtap_cilksort.cu:89-281. */
int gtap_cilksort(uint32_t* keys, uint32_t* tmp, int64_t n, int64_t leaf, int64_t merge_leaf,
                  gtap_run_stats* stats);
/* gtap_cilksort without a tmp argument and without allocating anything: the merge buffer is
 * the library's static device workspace (one per GPU, held until the process ends). No
 * counterpart in the reference: for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). Limit: n <= 4194304 keys (16 MiB);
 * larger n is refused with GTAP_ERROR_INVALID_ARGUMENT, and a call before gtap_initialize with
 * GTAP_ERROR_NOT_INITIALIZED. Calls that overlap on one GPU share the workspace and race. */
int gtap_cilksort_noalloc(uint32_t* keys, int64_t n, int64_t leaf, int64_t merge_leaf, gtap_run_stats* stats);
/* The same with a host buffer (copied in and out around the launch). */
int gtap_cilksort_host(uint32_t* keys, int64_t n, int64_t leaf, int64_t merge_leaf, gtap_run_stats* stats);

/* ------------------------------------------------------------ job streams */
/* A mixed stream of fib / n-queens / mergesort jobs run in one resident persistent kernel
 * (batched, or served while a host loop streams jobs in) or one launch per job (the plain
 * form). No reference counterpart: the reference runs one program per launch
 * (examples/fib/fib.cu:26-57). Kinds: 0 fib(n), 1 n-queens(N), 2 mergesort of `param` keys. */
typedef struct gtap_stream gtap_stream;
/* Job i of the stream from `seed`: 40% fib n in [18,28], 30% n-queens N in [8,12], 30%
 * mergesort of floor(2^(10+10u)) uint32 keys from the per-job key seed. */
int gtap_jobs_generate(uint64_t seed, int32_t njobs, int32_t* kind, int32_t* param, uint64_t* key_seed);
/* Generates the jobs, their sort inputs (device resident) and the job table. */
int gtap_stream_create(uint64_t seed, int32_t njobs, gtap_stream** out);
int gtap_stream_destroy(gtap_stream* stream);
int gtap_stream_info(gtap_stream* stream, int32_t* kind, int32_t* param, uint64_t* key_seed, int64_t* total_keys);
int gtap_stream_sorted_keys(gtap_stream* stream, int32_t job, uint32_t* out);
/* All jobs in one launch; latency_ms[i] = completion of job i after the launch start. */
int gtap_stream_run_batched(gtap_stream* stream, int64_t* results, double* latency_ms, gtap_run_stats* stats);
/* One launch per job, in order; latency_ms[i] = end of job i after the start of the sequence. */
int gtap_stream_run_per_job(gtap_stream* stream, int64_t* results, double* latency_ms, double* total_ms);
/* Poisson arrivals at `rate` jobs/s: served by one resident launch (per_job_launch = 0, at
 * most max_live jobs running at once, 0 -> 1024) or one launch per job (per_job_launch = 1).
 * latency_ms[i] = completion seen by the host - arrival. */
int gtap_stream_serve(gtap_stream* stream, double rate, uint64_t arrival_seed, int32_t per_job_launch,
                      int32_t max_live, int64_t* results, double* latency_ms, double* total_ms,
                      gtap_run_stats* stats);
/* Device time from admission to completion of each job of the last run (any of the three
 * forms above), in ms: the execution part of latency_ms. */
int gtap_stream_exec_ms(gtap_stream* stream, double* exec_ms);
/* Sort jobs merge runs of at most merge_leaf keys (512..16384, default 4096) on one wave and
 * split larger merges into merge tasks; the counterpart of TASK_SPAWN_CUTOFF_MERGE
 * (evaluation/2-comparison/cilksort/gtap_cilksort.cu:12, 112), as a run-time setting for the
 * autotuner. Results do not depend on it. */
int gtap_stream_set_merge_leaf(gtap_stream* stream, int32_t merge_leaf);

/* n-queens at the reference's benchmark settings (evaluation/2-comparison/nq/gtap_nq.cu:10-99):
 * tasks for rows 0..cutoff (the reference uses 7), a sequential search below, detached
 * spawns, one int32 counter. */
int gtap_nqueens_cutoff(int n, int cutoff, int64_t* result, gtap_run_stats* stats);

/* Block-mode BFS (examples/asynchronous_bfs/asynchronous_bfs_block.cu:17-37): one vertex
 * task per 256-thread workgroup, a thread per edge of ranges split to <= 128 edges. Same
 * buffers and result as gtap_bfs. */
int gtap_bfs_block(const int32_t* row, const int32_t* col, int32_t nv, int64_t ne, int32_t source, int32_t* depth,
                   gtap_run_stats* stats);

/* ------------------------------------------------------------ more reference workloads */
/* y = A x, CSR float64, divide-and-conquer over row ranges of at most `cutoff` rows (0 ->
 * 16), one task per row, detached spawns (examples/spmv/spmv_thread.cu:34-66). Device
 * buffers; the row sums are fma chains in column order. */
int gtap_spmv(const int32_t* row_ptr, const int32_t* col_idx, const double* val, const double* x, double* y,
              int32_t rows, int32_t cols, int32_t cutoff, gtap_run_stats* stats);
/* The same on host buffers (copied in and out). */
int gtap_spmv_host(const int32_t* row_ptr, const int32_t* col_idx, const double* val, const double* x, double* y,
                   int32_t rows, int32_t cols, int32_t cutoff, gtap_run_stats* stats);
/* A rows x rows random CSR matrix shaped like examples/spmv/spmv_thread.cu:89-107 (1 to
 * 2*avg_nnz-1 column draws per row, sorted, deduplicated; values in [-1, 1)). Call once with
 * col_idx = val = NULL for row_ptr and *nnz_out, then with buffers of that size. */
int gtap_spmv_random(int32_t rows, int32_t avg_nnz, uint64_t seed, int32_t* row_ptr, int32_t* col_idx, double* val,
                     int64_t* nnz_out);
/* The binary-tree workload (examples/binary_tree/binary_tree_thread.cu:63-120): a node task
 * spawns two children, waits, then reads mem_ops random input entries and runs
 * compute_iters fma steps; out[node] receives its float64 result. Device buffers; input_n a
 * power of two; 2^(height+1) - 1 nodes. */
int gtap_tree(int32_t height, int32_t mem_ops, int32_t compute_iters, const double* input, int64_t input_n,
              double* out, gtap_run_stats* stats);
/* The tree's input: n float64 in [-1, 1) from `seed` (splitmix64). */
int gtap_tree_input(uint64_t seed, int64_t n, double* out);

/* ------------------------------------------------------------ the reference's file formats */
/* Graph CSR (examples/asynchronous_bfs/asynchronous_bfs_thread.cu:69-84): int32 N, int32 M,
 * int32 row[N+1], int32 col[M]. */
int gtap_csr_file_info(const char* path, int32_t* nv, int64_t* ne);
int gtap_csr_file_read(const char* path, int32_t* row, int32_t* col);
int gtap_csr_file_write(const char* path, int32_t nv, int64_t ne, const int32_t* row, const int32_t* col);
/* Key vector (evaluation/util/gen_vector.c:21-60): size_t n, int32 data[n] (read as uint32
 * bit patterns). */
int gtap_keys_file_info(const char* path, uint64_t* n);
int gtap_keys_file_read(const char* path, uint32_t* keys);
int gtap_keys_file_write(const char* path, uint64_t n, const uint32_t* keys);
/* SpMV CSR (examples/spmv/spmv_thread.cu:70-86): int32 rows, cols, nnz, int32
 * row_ptr[rows+1], int32 col_idx[nnz], float64 val[nnz]. */
int gtap_spmv_file_info(const char* path, int32_t* rows, int32_t* cols, int32_t* nnz);
int gtap_spmv_file_read(const char* path, int32_t* row_ptr, int32_t* col_idx, double* val);
int gtap_spmv_file_write(const char* path, int32_t rows, int32_t cols, int32_t nnz, const int32_t* row_ptr,
                         const int32_t* col_idx, const double* val);

/* ------------------------------------------------------------ input generators (host) */
/* R-MAT edges, CSR by stable counting sort, and the gen_vector key stream
 * (evaluation/util/gen_vector.c:8-18; read format examples/asynchronous_bfs/
 * asynchronous_bfs_thread.cu:69-84). */
int gtap_rmat_edges(int scale, int64_t nedges, uint64_t seed, uint32_t ta, uint32_t tab, uint32_t tabc,
                    int32_t* src, int32_t* dst);
int gtap_csr_from_edges(int32_t nv, int64_t ne, const int32_t* src, const int32_t* dst, int32_t* row,
                        int32_t* col);
int gtap_rmat_csr(int scale, int edge_factor, uint64_t seed, uint32_t ta, uint32_t tab, uint32_t tabc,
                  int32_t* row, int32_t* col);
int gtap_gen_keys(uint64_t seed, uint64_t n, uint32_t* out);

#ifdef __cplusplus
}
/* The reference's unprefixed C++ spellings, so a host program that calls them compiles as is
 * (examples/cilksort/cilksort.cu:262). */
static inline void visualize_having_task_time(const char* app_name) { gtap_visualize_having_task_time(app_name); }
static inline void visualize_working_time(const char* app_name) { gtap_visualize_working_time(app_name); }
#endif
#endif
