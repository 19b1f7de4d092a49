// gtap_host.hip -- host runtime behind the C ABI (include/gtap.h).
//
// Reference host API: gtap_initialize / gtap_finalize / gtap_reset
// (runtime/detail/thread/gtap_runtime_thread.cuh:338-441), gtap_synchronize and the error
// report helpers (runtime/detail/common/gtap_runtime_error.cuh:38-249). Same names,
// int (hipError_t-compatible) status returns, no exceptions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../gtap_types.h"
#include "gtap_launch.hpp"

namespace gtap {
namespace host {

static Runtime g_runtime;
Runtime& rt() { return g_runtime; }

static const char* const kErrorShort[GTAP_ERROR_COUNT] = {
    "No error",
    "Invalid queue index",
    "Invalid queue index after join",
    "Queue overflow",
    "Result handle table overflow",
    "Task ID pool exhausted",
    "Task ID pool exhausted",
    "Generated task ID buffer overflow",
    "Invalid taskwait lowering",
    "Watchdog expired",
    "Invalid task function",
    "Invalid argument",
    "Runtime not initialized",
    "HIP error",
};

int fail(int code, const char* msg) {
    rt().last_error_code = code;
    fprintf(stderr, "GTaP error: %s (%s)\n", msg, (code >= 0 && code < GTAP_ERROR_COUNT) ? kErrorShort[code] : "?");
    return (int)hipErrorLaunchFailure;
}

int invalid_argument(const char* msg) {
    rt().last_error_code = GTAP_ERROR_INVALID_ARGUMENT;
    fprintf(stderr, "GTaP: invalid argument: %s\n", msg);
    return (int)hipErrorInvalidValue;
}

int check_csr_rows(const int32_t* row, int64_t n, int64_t ne, const char* who) {
    char msg[160];
    if (row == nullptr || n < 0 || ne < 0) {
        snprintf(msg, sizeof(msg), "%s: bad CSR arguments", who);
        return invalid_argument(msg);
    }
    if (row[0] != 0) {
        snprintf(msg, sizeof(msg), "%s: row offsets must start at 0 (found %d)", who, row[0]);
        return invalid_argument(msg);
    }
    for (int64_t i = 0; i < n; ++i) {
        if (row[i + 1] < row[i]) {
            snprintf(msg, sizeof(msg), "%s: row offsets decrease at row %lld", who, (long long)i);
            return invalid_argument(msg);
        }
    }
    if ((int64_t)row[n] != ne) {
        snprintf(msg, sizeof(msg), "%s: last row offset %d differs from the entry count %lld", who, row[n],
                 (long long)ne);
        return invalid_argument(msg);
    }
    return 0;
}

int hip_status(hipError_t e, const char* what) {
    fprintf(stderr, "GTaP HIP error: %s in %s\n", hipGetErrorString(e), what);
    return (int)e;
}

gtap_job make_job(uint32_t func, uint32_t queue, const void* args, size_t nbytes) {
    gtap_job j;
    memset(&j, 0, sizeof(j));
    j.func = func;
    j.queue = queue;
    j.status = GTAP_JOB_READY;
    if (nbytes > sizeof(j.args)) nbytes = sizeof(j.args);
    if (args && nbytes) memcpy(j.args, args, nbytes);
    return j;
}

static uint32_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return (uint32_t)p;
}

// Ids each worker holds at launch start: a wave's refill target and more than a workgroup's
// per-invocation spawn reserve.
constexpr uint32_t kStaticIds = 192;

GtapDev make_dev(uint32_t nworkers, uint32_t term_mode, uint32_t njobs, gtap_job* jobs, uint32_t job_ring) {
    Runtime& R = rt();
    GtapDev d;
    memset(&d, 0, sizeof(d));
    d.recs = R.recs;
    d.priv_ids = R.priv_ids;
    d.nshards = R.nshards;
    d.pool_ids = R.pool_ids;
    d.pool_seq = R.pool_seq;
    d.pool_ctl = R.pool_ctl;
    d.dq_word = R.dq_word;
    d.dq_slots = R.dq_slots;
    d.ctl = R.ctl;
    d.host_ctl = R.host_ctl_dev;
    d.report = R.report_dev;
    d.jobs = jobs;
    d.stats = R.stats;
    d.timeout_ticks = (uint64_t)R.cfg.timeout_ms * 100000ull;  // 100 MHz wall clock
    d.stride = R.stride;
    d.nrecs = R.nrecs;
    d.nanchors = R.nanchors;
    d.ncells = R.ncells;
    d.cap = R.cap;
    d.nqueues = R.nqueues;
    d.nworkers = nworkers;
    d.steal_batch = R.cfg.steal_batch;
    d.probe_lanes = R.cfg.probe_lanes;
    d.term_mode = term_mode;
    d.njobs = njobs;
    d.job_ring = job_ring;
    d.admit_budget = 0;
    d.priv_cap = kPrivCap;
    d.job_units = nullptr;
    d.max_live_jobs = R.nanchors;
    // A launch's first jobs go to its first workers by index, and every worker starts with a
    // fixed range of ids: starting touches neither the admission counter nor the id pool's
    // shared counter, which thousands of workers would otherwise all hit at once.
    if (term_mode != GTAP_TERM_SERVE && jobs != nullptr)
        d.static_jobs = std::min(std::min(njobs, nworkers), R.nanchors);
    if (64ull + (uint64_t)nworkers * kStaticIds <= R.nrecs / 2) d.static_ids = kStaticIds;
    d.prof_max = R.prof_max;
    d.gq_cap = R.gq_cap;
    d.gq_slots = R.gq_slots;
    d.gq_ctl = R.gq_ctl;
    d.ovf_cells = R.ovf_cells;
    d.ovf_ctl = R.ovf_ctl;
    d.ovf_cap = R.ovf_cap;
    d.ovf_epoch = R.ovf_epoch;
    d.epaq_publish = R.cfg.epaq_publish ? 1u : 0u;
    d.xcd_steal = R.cfg.xcd_steal ? 1u : 0u;
    d.idle_level = 1;
    for (uint32_t v = R.cfg.idle_sleep; v > 1 && d.idle_level < 6; v >>= 1) d.idle_level += 1;
    d.prof_having = R.prof_having;
    d.prof_phase = R.prof_phase;
    d.prof_working = R.prof_working;
    d.prof_counts = R.prof_counts;
    return d;
}

// Clears the per-launch state: control words, deque words, id-pool ring, statistics.
__global__ void reset_kernel(GtapDev dev, gtap_job job0, int32_t write_job0, int32_t arm) {
    const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t gs = (uint64_t)gridDim.x * blockDim.x;
    if (write_job0 && gid < sizeof(gtap_job) / 4)
        reinterpret_cast<uint32_t*>(dev.jobs)[gid] = reinterpret_cast<const uint32_t*>(&job0)[gid];
    // The workers that start a job by index count as started from the outset (quiescence
    // compares the idle count with the started count): a worker that owns a job but is not
    // resident yet must keep every other worker from taking the launch as finished.
    for (uint64_t i = gid; i < (uint64_t)CTL_COUNT * 32; i += gs)
        dev.ctl[i] = (i == CTL_JOB_NEXT * 32 || i == CTL_LIVE_JOBS * 32 || i == CTL_STARTED * 32)
                         ? (int32_t)dev.static_jobs
                         : (i == CTL_ARMED * 32 ? arm : 0);
    for (uint64_t i = gid; i < (uint64_t)dev.nqueues * dev.nworkers; i += gs) dev.dq_word[i * 16] = 0;
    for (uint64_t i = gid; i < (uint64_t)dev.nshards * dev.ncells; i += gs)
        dev.pool_seq[i] = (uint32_t)(i & (dev.ncells - 1));
    for (uint64_t i = gid; i < (uint64_t)dev.nworkers * ST_COUNT; i += gs) dev.stats[i] = 0;
    if (dev.gq_ctl)
        for (uint64_t i = gid; i < (uint64_t)dev.nqueues * 3 * 16; i += gs) dev.gq_ctl[i] = 0;
    if (dev.prof_phase)
        for (uint64_t i = gid; i < (uint64_t)dev.nworkers * PH_COUNT; i += gs) dev.prof_phase[i] = 0;
    if (dev.prof_max) {
        for (uint64_t i = gid; i < (uint64_t)dev.nworkers * dev.prof_max; i += gs) {
            dev.prof_having[i] = 0;
            dev.prof_working[i] = 0;
            dev.prof_counts[i] = 0;
        }
    }
    // ring positions start at 0; the fresh-id counter starts at 64 (ids 0..63 are
    // reserved, id 0 is the sink record), above the workers' fixed first ranges
    for (uint64_t i = gid; i < (uint64_t)dev.nshards * 32 + 32; i += gs)
        dev.pool_ctl[i] = (i == (uint64_t)dev.nshards * 32) ? 64ull + (uint64_t)dev.nworkers * dev.static_ids : 0ull;
}

// The overflow rings keep their positions from launch to launch. A launch that ended early
// (an error, first-task termination with detached tasks still queued) can leave tickets
// between the head and the ticket counter: their cells are freed for the next lap and the
// head moves up to the tickets. One workgroup, so the head moves after every cell is freed.
__global__ void __launch_bounds__(256) ovf_drain_kernel(GtapDev dev) {
    for (uint32_t q = 0; q < dev.nqueues; ++q) {
        const uint64_t head = ld64(dev.ovf_ctl + (size_t)q * 32), tickets = ld64(dev.ovf_ctl + (size_t)q * 32 + 16);
        for (uint64_t p = head + threadIdx.x; p < tickets; p += blockDim.x)
            st64(dev.ovf_cells + (size_t)q * dev.ovf_cap + (p & (dev.ovf_cap - 1)), ((p + dev.ovf_cap) & 0x3FFFFFFFull) << 32);
        __syncthreads();
        if (threadIdx.x == 0) st64(dev.ovf_ctl + (size_t)q * 32, tickets);
        __syncthreads();
    }
}

// Every cell free for its first lap (the tag of cell i is i), positions at 0.
__global__ void ovf_init_kernel(uint64_t* cells, uint64_t* ctl, uint32_t nqueues, uint32_t cap) {
    const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t gs = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = gid; i < (uint64_t)nqueues * cap; i += gs) cells[i] = (uint64_t)(i & (cap - 1)) << 32;
    for (uint64_t i = gid; i < (uint64_t)nqueues * 32; i += gs) ctl[i] = 0;
}

int enqueue_reset(const GtapDev& dev, hipStream_t stream, const gtap_job* job0, bool arm) {
    gtap_job j;
    if (job0) j = *job0;
    else memset(&j, 0, sizeof(j));
    if (dev.ovf_cells) {
        hipLaunchKernelGGL(ovf_drain_kernel, dim3(1), dim3(256), 0, stream, dev);
        GTAP_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(reset_kernel, dim3(1024), dim3(256), 0, stream, dev, j, job0 ? 1 : 0, arm ? 1 : 0);
    GTAP_HIP_TRY(hipGetLastError());
    return 0;
}

int reset_launch_state(const GtapDev& dev, hipStream_t stream, const gtap_job* job0, bool arm) {
    if (int st = enqueue_reset(dev, stream, job0, arm)) return st;
    Runtime& R = rt();
    if (R.report) memset(R.report, 0, sizeof(gtap_error_report));
    R.stats_host_valid = false;
    return 0;
}

// One workgroup: sums the per-worker statistics of the launch that just ended and writes
// them, the first error code and (on request) job-table slot 0 to the pinned summary.
constexpr int kSummaryThreads = 256;
__global__ void __launch_bounds__(kSummaryThreads) summary_kernel(GtapDev dev, LaunchSummary* out, int32_t copy_job0) {
    __shared__ uint64_t part[ST_COUNT][kSummaryThreads];
    const uint32_t t = threadIdx.x;
    uint64_t acc[ST_COUNT];
#pragma unroll
    for (int i = 0; i < ST_COUNT; ++i) acc[i] = 0;
    for (uint32_t w = t; w < dev.nworkers; w += kSummaryThreads) {
#pragma unroll
        for (int i = 0; i < ST_COUNT; ++i) acc[i] += dev.stats[(size_t)w * ST_COUNT + i];
    }
#pragma unroll
    for (int i = 0; i < ST_COUNT; ++i) part[i][t] = acc[i];
    __syncthreads();
    for (uint32_t h = kSummaryThreads / 2; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int i = 0; i < ST_COUNT; ++i) part[i][t] += part[i][t + h];
        }
        __syncthreads();
    }
    if (t < ST_COUNT) out->sums[t] = part[t][0];
    if (t == 0) {
        out->error_code = dev.ctl[CTL_ERROR * 32];
        out->admission_holds = dev.ctl[CTL_ADMIT_HOLDS * 32];
        out->has_job0 = copy_job0;
    }
    if (copy_job0 && t < sizeof(gtap_job) / 4)
        reinterpret_cast<uint32_t*>(&out->job0)[t] = reinterpret_cast<const uint32_t*>(dev.jobs)[t];
}

int enqueue_summary(const GtapDev& dev, hipStream_t stream, bool copy_job0) {
    Runtime& R = rt();
    hipLaunchKernelGGL(summary_kernel, dim3(1), dim3(kSummaryThreads), 0, stream, dev, R.summary_dev,
                       copy_job0 ? 1 : 0);
    GTAP_HIP_TRY(hipGetLastError());
    return 0;
}

constexpr uint32_t kGlobalQueueSlots = 1u << 24;  // per EPAQ queue
constexpr uint32_t kOverflowCells = 1u << 22;     // deque overflow ring cells per EPAQ queue (32 MiB)

int ensure_global_queues() {
    Runtime& R = rt();
    if (R.gq_slots) return 0;
    GTAP_HIP_TRY(hipMalloc(&R.gq_slots, (size_t)R.nqueues * kGlobalQueueSlots * sizeof(uint32_t)));
    GTAP_HIP_TRY(hipMalloc(&R.gq_ctl, (size_t)R.nqueues * 3 * 16 * sizeof(uint64_t)));
    GTAP_HIP_TRY(hipMemset(R.gq_ctl, 0, (size_t)R.nqueues * 3 * 16 * sizeof(uint64_t)));
    R.gq_cap = kGlobalQueueSlots;
    return 0;
}

int fetch_worker_stats() {
    Runtime& R = rt();
    if (R.stats_host_valid) return 0;
    const size_t nst = (size_t)R.last_nworkers * ST_COUNT;
    R.stats_host.resize(nst);
    if (nst) GTAP_HIP_TRY(hipMemcpy(R.stats_host.data(), R.stats, nst * sizeof(uint64_t), hipMemcpyDeviceToHost));
    R.stats_host_valid = true;
    return 0;
}

int occupancy_blocks(const void* kernel, uint32_t requested, uint32_t* blocks_out) {
    Runtime& R = rt();
    int per_cu = 0;
    GTAP_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0));
    if (per_cu < 1) per_cu = 1;
    uint32_t blocks = (uint32_t)per_cu * (uint32_t)R.num_cus;
    if (requested && requested < blocks) blocks = requested;
    *blocks_out = blocks;
    return 0;
}

int finish_launch(const GtapDev& dev, uint32_t blocks, gtap_run_stats* stats) {
    Runtime& R = rt();
    float ms = 0.f;
    GTAP_HIP_TRY(hipEventElapsedTime(&ms, R.ev0, R.ev1));
    const volatile LaunchSummary* sm = R.summary;
    const int32_t err = sm->error_code;
    gtap_run_stats s;
    memset(&s, 0, sizeof(s));
    s.kernel_ms = ms;
    s.tasks_executed = sm->sums[ST_TASKS];
    s.tasks_spawned = sm->sums[ST_SPAWNS];
    s.rounds = sm->sums[ST_ROUNDS];
    s.pops = sm->sums[ST_POPS];
    s.steals = sm->sums[ST_STEALS];
    s.steal_attempts = sm->sums[ST_STEAL_ATTEMPTS];
    s.jobs_taken = sm->sums[ST_JOBS];
    s.pool_refills = sm->sums[ST_REFILLS];
    s.published = sm->sums[ST_PUBLISHED];
    s.overflowed = sm->sums[ST_OVERFLOWED];
    s.suspended = sm->sums[ST_SUSPENDED];
    s.workers = dev.nworkers;
    s.blocks = blocks;
    R.last_nworkers = dev.nworkers;
    s.error_code = err;
    s.admission_holds = sm->admission_holds;
    R.last = s;
    R.last_error_code = err;
    if (stats) *stats = s;
    if (err != 0) {
        if (R.report && R.report->valid) {
            fprintf(stderr, "GTaP Runtime Error at block %d, thread %d: %s (code=%d, tid=%d, queue=%d, value=%d, "
                            "limit=%d) (source_line: %d)\n",
                    R.report->block_idx, R.report->thread_idx,
                    (err > 0 && err < GTAP_ERROR_COUNT) ? kErrorShort[err] : "?", R.report->code, R.report->tid,
                    R.report->queue_idx, R.report->value, R.report->limit, R.report->src_line);
        }
        return (int)hipErrorLaunchFailure;
    }
    return 0;
}

static void free_all() {
    Runtime& R = rt();
    auto dfree = [](void* p) {
        if (p) (void)hipFree(p);
    };
    dfree(R.recs);
    dfree(R.priv_ids);
    dfree(R.pool_ids);
    dfree(R.pool_seq);
    dfree(R.pool_ctl);
    dfree(R.dq_word);
    dfree(R.dq_slots);
    dfree(R.ctl);
    dfree(R.jobs);
    dfree(R.stats);
    dfree(R.prof_having);
    dfree(R.prof_phase);
    dfree(R.prof_working);
    dfree(R.prof_counts);
    dfree(R.gq_slots);
    dfree(R.gq_ctl);
    dfree(R.ovf_cells);
    dfree(R.ovf_ctl);
    dfree(R.ovf_epoch);
    dfree(R.job_units);
    if (R.host_ctl) (void)hipHostFree(R.host_ctl);
    if (R.report) (void)hipHostFree(R.report);
    if (R.summary) (void)hipHostFree(R.summary);
    if (R.jobs_host) (void)hipHostFree(R.jobs_host);
    if (R.ev0) (void)hipEventDestroy(R.ev0);
    if (R.ev1) (void)hipEventDestroy(R.ev1);
    if (R.stream) (void)hipStreamDestroy(R.stream);
    gtap_config keep = R.cfg;
    R = Runtime();
    R.cfg = keep;
}

// Lowered modules whose entry kernels the host launches (gtap_register_entry_module).
static std::vector<gtap_entry_setup_fn>& entry_modules() {
    static std::vector<gtap_entry_setup_fn> mods;
    return mods;
}

GtapDev entry_dev() {
    Runtime& R = rt();
    return make_dev(R.max_workers, GTAP_TERM_JOBS, 1, R.jobs, 1);
}

int arm_entry_modules() {
    Runtime& R = rt();
    const GtapDev d = entry_dev();
    if (int st = reset_launch_state(d, R.stream, nullptr, true)) return st;
    GTAP_HIP_TRY(hipStreamSynchronize(R.stream));
    // the per-worker counters gtap_get_worker_stats reads next are those of the entry launch
    // (the reset zeroed every provisioned worker's row; the launch fills the ones it runs)
    R.last_nworkers = d.nworkers;
    R.stats_host_valid = false;
    for (gtap_entry_setup_fn f : entry_modules())
        if (int st = f(&d, sizeof(d))) return hip_status((hipError_t)st, "lowered module setup");
    return 0;
}

// Largest frame over the programs compiled into this library, header included
// (the counterpart of the compiler-defined __gtap_auto_task_data_size).
uint32_t builtin_record_stride() { return kBuiltinStride; }

}  // namespace host
}  // namespace gtap

using namespace gtap::host;
using namespace gtap;

static void apply_defaults(gtap_config* c) {
    if (!c->pool_bytes) c->pool_bytes = 4ull << 30;
    if (!c->record_stride) c->record_stride = builtin_record_stride();
    if (!c->deque_capacity) c->deque_capacity = 32768;
    if (!c->num_queues) c->num_queues = 4;
    if (!c->steal_batch || c->steal_batch > 64) c->steal_batch = 64;
    if (!c->probe_lanes || c->probe_lanes > 64) c->probe_lanes = 8;
    if (!c->max_jobs) c->max_jobs = 65536;
    if (!c->timeout_ms) c->timeout_ms = 20000;
    if (!c->idle_sleep || c->idle_sleep > 32) c->idle_sleep = 32;
    if (c->scheduler == GTAP_SCHED_DEFAULT) c->scheduler = GTAP_SCHED_WORK_STEALING;
}

extern "C" {

int gtap_set_config(const gtap_config* cfg) {
    Runtime& R = rt();
    if (R.initialized) return invalid_argument("gtap_set_config must precede gtap_initialize");
    if (cfg && cfg->scheduler > GTAP_SCHED_CHASE_LEV) return invalid_argument("gtap_config.scheduler must be 0..3");
    if (cfg && cfg->epaq_publish > 1) return invalid_argument("gtap_config.epaq_publish must be 0 or 1");
    if (cfg && cfg->xcd_steal > 1) return invalid_argument("gtap_config.xcd_steal must be 0 or 1");
    if (cfg) R.cfg = *cfg;
    else memset(&R.cfg, 0, sizeof(R.cfg));
    return 0;
}

int gtap_get_config(gtap_config* out) {
    if (!out) return invalid_argument("null config");
    gtap_config c = rt().cfg;
    apply_defaults(&c);
    *out = c;
    return 0;
}

int gtap_initialize(void) {
    Runtime& R = rt();
    if (R.initialized) return 0;
    apply_defaults(&R.cfg);
    gtap_config& c = R.cfg;
    if (c.record_stride % 16 != 0 || c.record_stride < 32 || c.record_stride > 65520)
        return invalid_argument("record_stride must be a multiple of 16 in [32, 65520]");  // result offsets < 0xFFFF
    if ((c.deque_capacity & (c.deque_capacity - 1)) != 0 || c.deque_capacity < 4096)
        return invalid_argument("deque_capacity must be a power of two >= 4096");
    if (c.num_queues < 1 || c.num_queues > 8) return invalid_argument("num_queues must be in [1, 8]");
    GTAP_HIP_TRY(hipGetDevice(&R.device));
    hipDeviceProp_t prop;
    GTAP_HIP_TRY(hipGetDeviceProperties(&prop, R.device));
    R.num_cus = prop.multiProcessorCount;
    if (int st = cilksort_scratch_setup(R.device)) return st;
    GTAP_HIP_TRY(hipStreamCreateWithFlags(&R.stream, hipStreamNonBlocking));
    GTAP_HIP_TRY(hipEventCreate(&R.ev0));
    GTAP_HIP_TRY(hipEventCreate(&R.ev1));

    R.stride = c.record_stride;
    R.nanchors = c.max_jobs;
    uint64_t n = c.pool_bytes / R.stride;
    n = (n / 64) * 64;
    if (n < 4096) n = 4096;
    if (n + R.nanchors > 0x7FFFFFFFull) n = ((0x7FFFFFFFull - R.nanchors) / 64) * 64;
    // task records are addressed by 32-bit byte offsets (16-byte buffer accesses)
    if (n * R.stride > 0xFFFFFFFFull) n = ((0xFFFFFFFFull / R.stride) / 64) * 64;
    R.nrecs = (uint32_t)n;
    R.recs_bytes = (uint64_t)(R.nrecs + R.nanchors) * R.stride;
    GTAP_HIP_TRY(hipMalloc(&R.recs, R.recs_bytes));
    GTAP_HIP_TRY(hipMemset(R.recs, 0, R.recs_bytes));

    R.max_workers = (uint32_t)R.num_cus * 32;  // 8 workgroups x 4 waves per CU at most
    R.nshards = 8;
    R.ncells = next_pow2(R.nrecs / 64 + 1);  // every ring can hold every id
    GTAP_HIP_TRY(hipMalloc(&R.priv_ids, (size_t)R.max_workers * kPrivCap * sizeof(uint32_t)));
    GTAP_HIP_TRY(hipMalloc(&R.pool_ids, (size_t)R.nshards * R.ncells * 64 * sizeof(uint32_t)));
    GTAP_HIP_TRY(hipMalloc(&R.pool_seq, (size_t)R.nshards * R.ncells * sizeof(uint32_t)));
    GTAP_HIP_TRY(hipMalloc(&R.pool_ctl, ((size_t)R.nshards * 32 + 64) * sizeof(uint64_t)));
    GTAP_HIP_TRY(hipMemset(R.pool_ctl, 0, ((size_t)R.nshards * 32 + 64) * sizeof(uint64_t)));

    R.cap = c.deque_capacity;
    R.nqueues = c.num_queues;
    GTAP_HIP_TRY(hipMalloc(&R.dq_word, (size_t)R.nqueues * R.max_workers * 16 * sizeof(uint64_t)));
    GTAP_HIP_TRY(hipMemset(R.dq_word, 0, (size_t)R.nqueues * R.max_workers * 16 * sizeof(uint64_t)));
    GTAP_HIP_TRY(hipMalloc(&R.dq_slots, (size_t)R.nqueues * R.max_workers * R.cap * sizeof(uint32_t)));
    R.ovf_cap = kOverflowCells;
    GTAP_HIP_TRY(hipMalloc(&R.ovf_cells, (size_t)R.nqueues * R.ovf_cap * sizeof(uint64_t)));
    GTAP_HIP_TRY(hipMalloc(&R.ovf_ctl, (size_t)R.nqueues * 32 * sizeof(uint64_t)));
    GTAP_HIP_TRY(hipMalloc(&R.ovf_epoch, 64 * 32 * sizeof(uint32_t)));
    GTAP_HIP_TRY(hipMemset(R.ovf_epoch, 0, 64 * 32 * sizeof(uint32_t)));
    hipLaunchKernelGGL(ovf_init_kernel, dim3(1024), dim3(256), 0, R.stream, R.ovf_cells, R.ovf_ctl, R.nqueues,
                       R.ovf_cap);
    GTAP_HIP_TRY(hipGetLastError());

    GTAP_HIP_TRY(hipMalloc(&R.ctl, (size_t)CTL_COUNT * 32 * sizeof(int32_t)));
    GTAP_HIP_TRY(hipMemset(R.ctl, 0, (size_t)CTL_COUNT * 32 * sizeof(int32_t)));
    GTAP_HIP_TRY(hipMalloc(&R.jobs, (size_t)R.nanchors * sizeof(gtap_job)));
    GTAP_HIP_TRY(hipMemset(R.jobs, 0, (size_t)R.nanchors * sizeof(gtap_job)));
    GTAP_HIP_TRY(hipMalloc(&R.job_units, (size_t)R.nanchors * sizeof(uint32_t)));
    GTAP_HIP_TRY(hipMalloc(&R.stats, (size_t)R.max_workers * ST_COUNT * sizeof(uint64_t)));

    GTAP_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&R.host_ctl), HCTL_COUNT * 32 * sizeof(int32_t),
                               hipHostMallocMapped | hipHostMallocCoherent));
    memset(R.host_ctl, 0, HCTL_COUNT * 32 * sizeof(int32_t));
    GTAP_HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&R.host_ctl_dev), R.host_ctl, 0));
    GTAP_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&R.report), sizeof(gtap_error_report),
                               hipHostMallocMapped | hipHostMallocCoherent));
    memset(R.report, 0, sizeof(gtap_error_report));
    GTAP_HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&R.report_dev), R.report, 0));
    GTAP_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&R.summary), sizeof(LaunchSummary),
                               hipHostMallocMapped | hipHostMallocCoherent));
    memset(R.summary, 0, sizeof(LaunchSummary));
    GTAP_HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&R.summary_dev), R.summary, 0));
    GTAP_HIP_TRY(hipDeviceSynchronize());
    R.initialized = true;
    // ready for one host-launched entry kernel, as the reference's state is after initialization
    return arm_entry_modules();
}

int gtap_finalize(void) {
    Runtime& R = rt();
    if (!R.initialized) return 0;
    (void)hipDeviceSynchronize();
    free_all();
    return 0;
}

// Clears the runtime state without reallocating (every launch also does this itself) and
// arms it for one host-launched entry kernel of a lowered module.
int gtap_reset(void) {
    Runtime& R = rt();
    if (!R.initialized) return fail(GTAP_ERROR_NOT_INITIALIZED, "gtap_reset before gtap_initialize");
    if (int st = arm_entry_modules()) return st;
    R.last_error_code = 0;
    return 0;
}

// Only records the module: the descriptor goes to it at the next gtap_initialize or gtap_reset.
// A module registers from its static initializer, which may run before the HIP runtime has
// registered that module's device symbols, so no symbol may be written from here.
int gtap_register_entry_module(gtap_entry_setup_fn setup) {
    if (setup == nullptr) return invalid_argument("gtap_register_entry_module: null setup function");
    std::vector<gtap_entry_setup_fn>& mods = entry_modules();
    if (std::find(mods.begin(), mods.end(), setup) == mods.end()) mods.push_back(setup);
    return 0;
}

int gtap_get_runtime_error_code(int* code) {
    Runtime& R = rt();
    if (!code) return invalid_argument("null pointer");
    if (!R.initialized) {
        *code = R.last_error_code;  // a host-side error (e.g. a refused configuration)
        return 0;
    }
    int32_t e = 0;
    GTAP_HIP_TRY(hipMemcpy(&e, R.ctl + CTL_ERROR * 32, sizeof(e), hipMemcpyDeviceToHost));
    *code = e ? e : R.last_error_code;
    return 0;
}

const char* gtap_get_runtime_error_string(int code) {
    if (code >= 0 && code < GTAP_ERROR_COUNT) return kErrorShort[code];
    return "Unknown error";
}

int gtap_print_runtime_error_report(void) {
    Runtime& R = rt();
    if (!R.report || !R.report->valid) return 0;
    const gtap_error_report* r = R.report;
    printf("GTaP Runtime Error at block %d, thread %d: %s (code=%d, tid=%d, queue=%d, value=%d, limit=%d) "
           "(source_line: %d)\n",
           r->block_idx, r->thread_idx, gtap_get_runtime_error_string(r->code), r->code, r->tid, r->queue_idx,
           r->value, r->limit, r->src_line);
    return 1;
}

int gtap_get_runtime_error_report(gtap_error_report* out) {
    Runtime& R = rt();
    if (!out) return invalid_argument("null pointer");
    if (R.report) *out = *R.report;
    else memset(out, 0, sizeof(*out));
    return 0;
}

int gtap_check_runtime_error(void) {
    if (gtap_print_runtime_error_report()) return 0;
    int code = 0;
    int st = gtap_get_runtime_error_code(&code);
    if (st) return st;
    if (code != GTAP_ERROR_NONE) printf("GTaP Runtime Error: %s (code: %d)\n", gtap_get_runtime_error_string(code), code);
    return 0;
}

int gtap_synchronize(void) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        if (!gtap_print_runtime_error_report()) printf("HIP ERROR: %s\n", hipGetErrorString(e));
        return (int)e;
    }
    return gtap_check_runtime_error();
}

// Timeline profiling for the following launches: up to max_samples stamps per worker of
// each kind (reference: PROFILE / MAX_PROFILE_DATA, runtime/detail/common/gtap_config.cuh:22-26).
int gtap_set_profiling(int enable, uint32_t max_samples) {
    Runtime& R = rt();
    if (!R.initialized) {
        int st = gtap_initialize();
        if (st) return st;
    }
    if (R.prof_having) (void)hipFree(R.prof_having);
    if (R.prof_working) (void)hipFree(R.prof_working);
    if (R.prof_counts) (void)hipFree(R.prof_counts);
    if (R.prof_phase) (void)hipFree(R.prof_phase);
    R.prof_having = R.prof_working = nullptr;
    R.prof_counts = nullptr;
    R.prof_phase = nullptr;
    R.prof_max = 0;
    if (!enable) return 0;
    if (max_samples == 0) max_samples = 4096;
    max_samples = (max_samples + 1u) & ~1u;  // start/end stamps come in pairs
    const size_t n = (size_t)R.max_workers * max_samples;
    GTAP_HIP_TRY(hipMalloc(&R.prof_having, n * sizeof(uint64_t)));
    GTAP_HIP_TRY(hipMalloc(&R.prof_working, n * sizeof(uint64_t)));
    GTAP_HIP_TRY(hipMalloc(&R.prof_counts, n * sizeof(int32_t)));
    GTAP_HIP_TRY(hipMalloc(&R.prof_phase, (size_t)R.max_workers * PH_COUNT * sizeof(uint64_t)));
    GTAP_HIP_TRY(hipMemset(R.prof_phase, 0, (size_t)R.max_workers * PH_COUNT * sizeof(uint64_t)));
    R.prof_max = max_samples;
    return 0;
}

// Shader-clock cycles of the last thread-mode launch's round phases, summed over its workers
// (out[PH_ACQUIRE .. PH_IDLE]), and its rounds (out[PH_ROUNDS]); needs timeline profiling on
// (gtap_set_profiling) for that launch, else everything reads 0.
int gtap_get_round_phases(uint64_t* out, uint32_t n) {
    Runtime& R = rt();
    if (!out || n < (uint32_t)PH_ROUNDS + 1) return invalid_argument("gtap_get_round_phases: out needs 9 entries");
    for (uint32_t k = 0; k < n; ++k) out[k] = 0;
    if (!R.initialized || !R.prof_phase || !R.last_nworkers) return 0;
    std::vector<uint64_t> h((size_t)R.last_nworkers * PH_COUNT);
    GTAP_HIP_TRY(hipMemcpy(h.data(), R.prof_phase, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    for (uint32_t w = 0; w < R.last_nworkers; ++w)
        for (int k = 0; k <= PH_ROUNDS; ++k) out[k] += h[(size_t)w * PH_COUNT + k];
    return 0;
}

int gtap_get_last_stats(gtap_run_stats* out) {
    if (!out) return invalid_argument("null pointer");
    *out = rt().last;
    return 0;
}

int gtap_get_worker_stats(uint64_t* out, uint32_t max_workers, uint32_t* nworkers_out) {
    Runtime& R = rt();
    if (R.initialized) {
        if (int st = fetch_worker_stats()) return st;
    }
    const uint32_t nw = (uint32_t)(R.stats_host.size() / ST_COUNT);
    if (nworkers_out) *nworkers_out = nw;
    if (out) {
        const uint32_t n = std::min(nw, max_workers);
        memcpy(out, R.stats_host.data(), (size_t)n * ST_COUNT * sizeof(uint64_t));
    }
    return 0;
}

int gtap_device_info(char* name, int name_len, int* num_cus, uint64_t* hbm_bytes) {
    int dev = 0;
    GTAP_HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    GTAP_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s", prop.gcnArchName);
    }
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (uint64_t)prop.totalGlobalMem;
    return 0;
}

// ---- device memory helpers for callers that hold only host pointers (ctypes, C)
int gtap_device_malloc(void** ptr, uint64_t bytes) {
    if (!ptr) return invalid_argument("null pointer");
    GTAP_HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}
int gtap_device_free(void* ptr) {
    if (ptr) GTAP_HIP_TRY(hipFree(ptr));
    return 0;
}
// The helpers complete before they return: the runtime launches on its own non-blocking
// stream, which does not wait for work on the null stream, so a copy still in flight would
// race with the next launch that reads its destination (a device-to-device copy and a memset
// may return before they are done).
int gtap_memcpy_htod(void* dst, const void* src, uint64_t bytes) {
    if (bytes) GTAP_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    GTAP_HIP_TRY(hipStreamSynchronize(nullptr));
    return 0;
}
int gtap_memcpy_dtoh(void* dst, const void* src, uint64_t bytes) {
    Runtime& R = rt();
    if (R.initialized) GTAP_HIP_TRY(hipStreamSynchronize(R.stream));  // the launches that wrote src
    if (bytes) GTAP_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}
int gtap_memcpy_dtod(void* dst, const void* src, uint64_t bytes) {
    Runtime& R = rt();
    if (R.initialized) GTAP_HIP_TRY(hipStreamSynchronize(R.stream));
    if (bytes) GTAP_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    GTAP_HIP_TRY(hipStreamSynchronize(nullptr));
    return 0;
}
int gtap_memset_device(void* dst, int value, uint64_t bytes) {
    Runtime& R = rt();
    if (R.initialized) GTAP_HIP_TRY(hipStreamSynchronize(R.stream));
    if (bytes) GTAP_HIP_TRY(hipMemset(dst, value, bytes));
    GTAP_HIP_TRY(hipStreamSynchronize(nullptr));
    return 0;
}

}  // extern "C"
