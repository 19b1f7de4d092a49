// gtap_launch.hpp -- host side of one persistent launch (templated on the task program).
//
// The host runtime owns every pool (task records, id pool, deques, control words, job
// table); a launch resets the per-launch state with one kernel, records hipEvents around
// the persistent kernel and reads back the statistics and any device error.
// Reference host flow: gtap_initialize / launch / synchronize
// (runtime/detail/thread/gtap_runtime_thread.cuh:50-280, examples/fib/fib.cu:26-57).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../device/gtap_block.hpp"
#include "../gtap_types.h"

extern "C" int gtap_initialize(void);
extern "C" int gtap_register_entry_module(gtap_entry_setup_fn setup);

namespace gtap {
namespace host {

// Record stride of the built-in programs: the largest frame plus the 16-byte header
// (n-queens: 16 args + 8 + 16 int64 result slots = 152 bytes of frame).
constexpr uint32_t kBuiltinStride = 192;

// Written by the summary kernel after each launch into pinned host memory: the per-worker
// statistics summed, the first error code, and job-table slot 0 (the root of a single-job
// launch). The host reads it after the stream synchronises, with no copy calls.
struct LaunchSummary {
    uint64_t sums[ST_COUNT];
    int32_t error_code;
    int32_t has_job0;
    int32_t admission_holds;
    int32_t pad;
    gtap_job job0;
};

struct Runtime {
    bool initialized = false;
    int device = 0;
    gtap_config cfg{};
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // task records and anchors
    uint8_t* recs = nullptr;
    uint64_t recs_bytes = 0;
    uint32_t stride = 0;
    uint32_t nrecs = 0;
    uint32_t nanchors = 0;
    // id pool
    uint32_t* priv_ids = nullptr;
    uint32_t nshards = 0;
    uint32_t* pool_ids = nullptr;
    uint32_t* pool_seq = nullptr;
    uint64_t* pool_ctl = nullptr;
    uint32_t ncells = 0;
    // deques
    uint64_t* dq_word = nullptr;
    uint32_t* dq_slots = nullptr;
    uint32_t cap = 0;
    uint32_t nqueues = 0;
    uint32_t max_workers = 0;
    // control, errors, jobs, statistics
    int32_t* ctl = nullptr;
    int32_t* host_ctl = nullptr;       // pinned, device-mapped
    int32_t* host_ctl_dev = nullptr;
    gtap_error_report* report = nullptr;      // pinned, device-mapped
    gtap_error_report* report_dev = nullptr;
    gtap_job* jobs = nullptr;          // device job table
    gtap_job* jobs_host = nullptr;     // pinned job table (served mode)
    gtap_job* jobs_host_dev = nullptr;
    uint32_t* job_units = nullptr;     // served: device copy of each job slot's reservation
    uint64_t* stats = nullptr;
    std::vector<uint64_t> stats_host;  // per-worker statistics of the last launch, fetched on demand
    bool stats_host_valid = false;
    LaunchSummary* summary = nullptr;  // pinned, device-mapped
    LaunchSummary* summary_dev = nullptr;
    // global-queue scheduler (allocated on first use)
    uint32_t* gq_slots = nullptr;
    uint64_t* gq_ctl = nullptr;
    // deque overflow rings (work stealing): [queue][ovf_cap] cells, head/ticket words
    uint64_t* ovf_cells = nullptr;
    uint64_t* ovf_ctl = nullptr;
    uint32_t* ovf_epoch = nullptr;
    uint32_t ovf_cap = 0;
    uint32_t gq_cap = 0;
    // timeline profiling (gtap_set_profiling)
    uint32_t prof_max = 0;
    uint64_t* prof_having = nullptr;
    uint64_t* prof_working = nullptr;
    int32_t* prof_counts = nullptr;
    uint64_t* prof_phase = nullptr;  // [worker][PH_COUNT] round-phase cycles (thread-mode timeline build)
    uint32_t last_nworkers = 0;
    gtap_run_stats last{};
    int last_error_code = 0;
};

Runtime& rt();
int fail(int code, const char* msg);
int invalid_argument(const char* msg);
int hip_status(hipError_t e, const char* what);
// Row offsets of a CSR structure in host memory: row[0] == 0, non-decreasing, row[n] == ne.
// Returns 0, or GTAP_ERROR_INVALID_ARGUMENT's status with `who` in the message.
int check_csr_rows(const int32_t* row, int64_t n, int64_t ne, const char* who);
// Resolves the device address of gtap_cilksort_noalloc's static workspace on the current
// device (prog_cilksort.hip); gtap_initialize calls it.
int cilksort_scratch_setup(int device);
gtap_job make_job(uint32_t func, uint32_t queue, const void* args, size_t nbytes);
GtapDev make_dev(uint32_t nworkers, uint32_t term_mode, uint32_t njobs, gtap_job* jobs, uint32_t job_ring);
// Resets the per-launch device state; a non-null job0 is written into job-table slot 0 by
// the reset kernel (a single-job launch needs no host-to-device copy).
// `arm` = 1 lets one host-launched entry kernel run on the state (CTL_ARMED).
int reset_launch_state(const GtapDev& dev, hipStream_t stream, const gtap_job* job0 = nullptr, bool arm = false);
// Enqueues the per-launch state reset kernel only (usable under stream capture).
int enqueue_reset(const GtapDev& dev, hipStream_t stream, const gtap_job* job0 = nullptr, bool arm = false);
// The descriptor of a host-launched entry kernel: every provisioned worker, one job (slot 0,
// started by worker 0), first-task termination (a module sets its own mode).
GtapDev entry_dev();
// Resets and arms the state for one host-launched entry kernel and pushes the descriptor to
// every registered lowered module (gtap_initialize, gtap_reset).
int arm_entry_modules();
// Enqueues the summary kernel (after the launch's end event): it fills R.summary.
int enqueue_summary(const GtapDev& dev, hipStream_t stream, bool copy_job0);
// After the stream has synchronised: kernel time, summed statistics and the error check.
int finish_launch(const GtapDev& dev, uint32_t blocks, gtap_run_stats* stats);
// The per-worker statistics of the last launch (copied from the device on first use).
int fetch_worker_stats();
int occupancy_blocks(const void* kernel, uint32_t requested, uint32_t* blocks_out);

#define GTAP_HIP_TRY(call)                                              \
    do {                                                                \
        hipError_t _e = (call);                                         \
        if (_e != hipSuccess) return ::gtap::host::hip_status(_e, #call); \
    } while (0)

// Persistent launch of a thread-mode program over `njobs` jobs of the job table `jobs`.
// Checks a thread-mode launch against the provisioned runtime and picks its workgroup count
// (every co-resident workgroup unless gtap_config.blocks asks for fewer).
template <class P, int Sched = kSchedWS, bool Prof = false>
int thread_launch_shape(uint32_t job_ring, uint32_t* blocks_out) {
    Runtime& R = rt();
    if (!R.initialized) {
        int st = gtap_initialize();
        if (st) return st;
    }
    if ((uint32_t)P::kQueues > R.nqueues) return invalid_argument("program needs more EPAQ queues than provisioned");
    if (P::kFrameBytes + kHdrBytes > R.stride) return invalid_argument("task frame exceeds the record stride");
    if (job_ring == 0 || job_ring > R.nanchors) return invalid_argument("job ring exceeds the anchor capacity");
    static uint32_t max_blocks = 0;
    const void* kfn = reinterpret_cast<const void*>(&thread_kernel<P, Sched, Prof>);
    if (max_blocks == 0) {
        int st = occupancy_blocks(kfn, 0, &max_blocks);
        if (st) return st;
    }
    uint32_t blocks = max_blocks;
    if (R.cfg.blocks && R.cfg.blocks < blocks) blocks = R.cfg.blocks;
    if (blocks * 4 > R.max_workers) blocks = R.max_workers / 4;
    *blocks_out = blocks;
    return 0;
}

// The ring of the global-queue scheduler, allocated on first use.
int ensure_global_queues();

// A program that is also built with the alternative schedulers (P::kSchedVariants).
template <class P, class = void>
struct HasSchedVariants {
    static constexpr bool value = false;
};
template <class P>
struct HasSchedVariants<P, decltype((void)P::kSchedVariants)> {
    static constexpr bool value = P::kSchedVariants;
};

template <class P, int Sched, bool Prof>
int launch_thread_prof(const typename P::Args& args, uint32_t term_mode, uint32_t njobs, gtap_job* jobs,
                       uint32_t job_ring, gtap_run_stats* stats, hipStream_t stream, const gtap_job* job0) {
    Runtime& R = rt();
    uint32_t blocks = 0;
    if (int st = thread_launch_shape<P, Sched, Prof>(job_ring, &blocks)) return st;
    if (Sched == kSchedGQ) {
        if (int st = ensure_global_queues()) return st;
    }
    if (!stream) stream = R.stream;
    GtapDev dev = make_dev(blocks * 4, term_mode, njobs, jobs, job_ring);
    int st = reset_launch_state(dev, stream, job0);
    if (st) return st;
    GTAP_HIP_TRY(hipEventRecord(R.ev0, stream));
    hipLaunchKernelGGL((thread_kernel<P, Sched, Prof>), dim3(blocks), dim3(256), 0, stream, dev, args);
    GTAP_HIP_TRY(hipGetLastError());
    GTAP_HIP_TRY(hipEventRecord(R.ev1, stream));
    if ((st = enqueue_summary(dev, stream, job0 != nullptr))) return st;
    GTAP_HIP_TRY(hipStreamSynchronize(stream));
    return finish_launch(dev, blocks, stats);
}

// The timeline-profiling build of the kernel when profiling is on, the plain one otherwise.
template <class P, int Sched>
int launch_thread_as(const typename P::Args& args, uint32_t term_mode, uint32_t njobs, gtap_job* jobs,
                     uint32_t job_ring, gtap_run_stats* stats, hipStream_t stream, const gtap_job* job0) {
    if (rt().prof_max) return launch_thread_prof<P, Sched, true>(args, term_mode, njobs, jobs, job_ring, stats, stream, job0);
    return launch_thread_prof<P, Sched, false>(args, term_mode, njobs, jobs, job_ring, stats, stream, job0);
}

// Launches with the scheduler gtap_config.scheduler names (batched work stealing unless a
// program built with the alternatives asks for one of them).
template <class P>
int launch_thread(const typename P::Args& args, uint32_t term_mode, uint32_t njobs, gtap_job* jobs,
                  uint32_t job_ring, gtap_run_stats* stats, hipStream_t stream = nullptr,
                  const gtap_job* job0 = nullptr) {
    const uint32_t sched = rt().cfg.scheduler;
    if constexpr (HasSchedVariants<P>::value) {
        if (sched == GTAP_SCHED_GLOBAL_QUEUE)
            return launch_thread_as<P, kSchedGQ>(args, term_mode, njobs, jobs, job_ring, stats, stream, job0);
        if (sched == GTAP_SCHED_CHASE_LEV)
            return launch_thread_as<P, kSchedCL>(args, term_mode, njobs, jobs, job_ring, stats, stream, job0);
    } else {
        if (sched > GTAP_SCHED_WORK_STEALING)
            return invalid_argument("this program is built for the work-stealing scheduler only");
    }
    return launch_thread_as<P, kSchedWS>(args, term_mode, njobs, jobs, job_ring, stats, stream, job0);
}

// A thread-mode launch held as an instantiated HIP graph: [copy of the job table from a
// device-resident template into the runtime's job table, state reset kernel, persistent
// kernel]. Replaying it costs one hipGraphLaunch. `dev` is the launch's argument block: the
// graph is valid while make_dev() returns the same block (same buffers, same configuration).
struct ThreadGraph {
    hipGraphExec_t exec = nullptr;
    GtapDev dev{};
    uint32_t blocks = 0;
};

inline void destroy_graph(ThreadGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
}

template <class P>
int capture_thread(ThreadGraph& g, const typename P::Args& args, uint32_t term_mode, uint32_t njobs,
                   const gtap_job* jobs_template_dev, uint32_t job_ring) {
    Runtime& R = rt();
    uint32_t blocks = 0;
    const bool prof = R.prof_max != 0;
    if (int st = prof ? thread_launch_shape<P, kSchedWS, true>(job_ring, &blocks)
                      : thread_launch_shape<P, kSchedWS, false>(job_ring, &blocks))
        return st;
    GtapDev dev = make_dev(blocks * 4, term_mode, njobs, R.jobs, job_ring);
    if (g.exec && g.blocks == blocks && memcmp(&g.dev, &dev, sizeof(GtapDev)) == 0) return 0;
    destroy_graph(g);
    hipStream_t s = R.stream;
    GTAP_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipMemcpyAsync(R.jobs, jobs_template_dev, (size_t)njobs * sizeof(gtap_job),
                                  hipMemcpyDeviceToDevice, s);
    int st = (e == hipSuccess) ? enqueue_reset(dev, s) : hip_status(e, "hipMemcpyAsync (capture)");
    if (st == 0) {
        if (prof) hipLaunchKernelGGL((thread_kernel<P, kSchedWS, true>), dim3(blocks), dim3(256), 0, s, dev, args);
        else hipLaunchKernelGGL((thread_kernel<P, kSchedWS, false>), dim3(blocks), dim3(256), 0, s, dev, args);
        e = hipGetLastError();
        if (e != hipSuccess) st = hip_status(e, "thread_kernel (capture)");
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(s, &graph);
    if (st) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return hip_status(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g.exec = nullptr;
        return hip_status(e, "hipGraphInstantiate");
    }
    g.dev = dev;
    g.blocks = blocks;
    return 0;
}

// Replays a captured launch; stats->kernel_ms covers the whole graph (copy, reset, kernel).
inline int replay_thread(const ThreadGraph& g, gtap_run_stats* stats) {
    Runtime& R = rt();
    if (!g.exec) return invalid_argument("replay_thread: no captured graph");
    if (R.report) memset(R.report, 0, sizeof(gtap_error_report));
    GTAP_HIP_TRY(hipEventRecord(R.ev0, R.stream));
    GTAP_HIP_TRY(hipGraphLaunch(g.exec, R.stream));
    GTAP_HIP_TRY(hipEventRecord(R.ev1, R.stream));
    if (int st = enqueue_summary(g.dev, R.stream, false)) return st;
    GTAP_HIP_TRY(hipStreamSynchronize(R.stream));
    return finish_launch(g.dev, g.blocks, stats);
}

// One job in slot 0 of the device job table; the root's return value goes to *result.
template <class P>
int run_single_job(const typename P::Args& args, const gtap_job& job, int64_t* result, gtap_run_stats* stats,
                   uint32_t term_mode, gtap_job* job_out = nullptr) {
    Runtime& R = rt();
    if (!R.initialized) {
        int st = gtap_initialize();
        if (st) return st;
    }
    int st = launch_thread<P>(args, term_mode, 1, R.jobs, 1, stats, nullptr, &job);
    if (st) return st;
    const gtap_job out = R.summary->job0;
    if (term_mode == GTAP_TERM_JOBS && out.status != GTAP_JOB_DONE) return fail(GTAP_ERROR_INVALID_TASKWAIT, "job did not complete");
    if (result) *result = out.result;
    if (job_out) *job_out = out;
    return 0;
}

// Persistent launch of a block-mode program: one worker (and one deque) per workgroup.
template <class P>
int launch_block(const typename P::Args& args, uint32_t term_mode, uint32_t njobs, gtap_job* jobs, uint32_t job_ring,
                 gtap_run_stats* stats, hipStream_t stream = nullptr, const gtap_job* job0 = nullptr) {
    Runtime& R = rt();
    if (!R.initialized) {
        int st = gtap_initialize();
        if (st) return st;
    }
    if (P::kFrameBytes + kHdrBytes > R.stride) return invalid_argument("task frame exceeds the record stride");
    if (job_ring == 0 || job_ring > R.nanchors) return invalid_argument("job ring exceeds the anchor capacity");
    if (!stream) stream = R.stream;
    static uint32_t max_blocks = 0;
    const void* kfn = reinterpret_cast<const void*>(&block_kernel<P>);
    if (max_blocks == 0) {
        int st = occupancy_blocks(kfn, 0, &max_blocks);
        if (st) return st;
    }
    uint32_t blocks = max_blocks;
    if (R.cfg.block_workers && R.cfg.block_workers < blocks) blocks = R.cfg.block_workers;
    if (blocks > R.max_workers) blocks = R.max_workers;
    GtapDev dev = make_dev(blocks, term_mode, njobs, jobs, job_ring);
    int st = reset_launch_state(dev, stream, job0);
    if (st) return st;
    GTAP_HIP_TRY(hipEventRecord(R.ev0, stream));
    hipLaunchKernelGGL(block_kernel<P>, dim3(blocks), dim3(kBlockThreads), 0, stream, dev, args);
    GTAP_HIP_TRY(hipGetLastError());
    GTAP_HIP_TRY(hipEventRecord(R.ev1, stream));
    if ((st = enqueue_summary(dev, stream, job0 != nullptr))) return st;
    GTAP_HIP_TRY(hipStreamSynchronize(stream));
    return finish_launch(dev, blocks, stats);
}

template <class P>
int run_single_block_job(const typename P::Args& args, const gtap_job& job, int64_t* result, gtap_run_stats* stats,
                         uint32_t term_mode = GTAP_TERM_JOBS) {
    Runtime& R = rt();
    if (!R.initialized) {
        int st = gtap_initialize();
        if (st) return st;
    }
    int st = launch_block<P>(args, term_mode, 1, R.jobs, 1, stats, nullptr, &job);
    if (st) return st;
    const gtap_job out = R.summary->job0;
    if (term_mode == GTAP_TERM_JOBS && out.status != GTAP_JOB_DONE) return fail(GTAP_ERROR_INVALID_TASKWAIT, "job did not complete");
    if (result) *result = out.result;
    return 0;
}

}  // namespace host
}  // namespace gtap
