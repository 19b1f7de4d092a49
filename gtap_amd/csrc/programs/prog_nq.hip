// prog_nq.hip -- n-queens as a thread-executed task program.
//
// Source program (the task shape of the reference's evaluation/2-comparison/nq/gtap_nq.cu:25-55,
// with no cutoff and with every count returned to the parent through a taskwait and an
// int64 sum instead of a global atomicAdd):
//
//   #pragma gtap function
//   __device__ int64_t nq(int row, uint32_t cols, uint32_t diag, uint32_t anti) {
//       if (row == N) return 1;
//       uint32_t avail = mask & ~(cols | diag | anti);
//       int64_t counts[16]; int k = 0;
//       while (avail) {
//           uint32_t p = avail & -avail; avail -= p;
//           #pragma gtap task
//           counts[k++] = nq(row + 1, cols | p, (diag | p) << 1, (anti | p) >> 1);
//       }
//       #pragma gtap taskwait
//       int64_t s = 0; for (int i = 0; i < k; ++i) s += counts[i];
//       return s;
//   }
//
// Two lowerings of this program:
//  * additive join (Program, the default): the counts are only summed after the taskwait,
//    so each child adds its count to the parent's join word with the atomic that counts the
//    join down; frame = { int32 row; uint32 cols, diag, anti; }, the parent resumes with
//    the sum.
//  * result slots (SlotProgram): frame = { int32 row; uint32 cols, diag, anti; int32 k;
//    int32 pad[3]; int64 counts[16]; }; each child stores its count into counts[i], then
//    counts the join down. The first six words (the arguments and k) are loaded when the
//    task starts, so the resumed task knows k without another load; counts[] is 16-byte
//    aligned and read two at a time.
#include "../device/gtap_abi.hpp"
#include "../host/gtap_launch.hpp"

namespace gtap {
namespace nq {

enum : uint32_t { F_NQ = 1 };
struct NqArg {
    int32_t row;
    uint32_t cols;
    uint32_t diag;
    uint32_t anti;
};
constexpr int kOffK = 16;
constexpr int kOffCounts = 32;
constexpr int kMaxN = 16;

template <bool kSum>
struct ProgramT {
    static constexpr int kQueues = 1;
    // the default form is also built with the global-queue and Chase-Lev schedulers
    static constexpr bool kSchedVariants = kSum;
    static constexpr int kWavesPerSimd = kSum ? 4 : 1;
    static constexpr int kFrameWords = kSum ? 4 : 6;  // row, cols, diag, anti (, k, pad)
    static constexpr uint32_t kMaxFunc = F_NQ;
    static constexpr uint32_t kFrameBytes = kSum ? 16 : kOffCounts + 8 * kMaxN;
    struct Args {
        int32_t n;
    };
    template <class L>
    __device__ __forceinline__ static void run(L& lane);
};
using Program = ProgramT<true>;
using SlotProgram = ProgramT<false>;

template <bool kSum, class L>
__device__ __forceinline__ void nq_task(L& lane) {
    const int32_t N = lane.args->n;
    switch (__gtap_get_task_state(lane)) {
        case 0: {
            const int32_t row = lane.template arg<int32_t>(0);
            const uint32_t cols = lane.template arg<uint32_t>(4);
            const uint32_t diag = lane.template arg<uint32_t>(8);
            const uint32_t anti = lane.template arg<uint32_t>(12);
            if (row == N) {
                __gtap_return(lane, (int64_t)1);
                return;
            }
            const uint32_t mask = (N < 32) ? ((1u << N) - 1u) : 0xFFFFFFFFu;
            uint32_t avail = mask & ~(cols | diag | anti);
            int32_t k = 0;
            while (avail) {
                const uint32_t p = avail & (0u - avail);
                avail -= p;
                const NqArg child{row + 1, cols | p, (diag | p) << 1, (anti | p) >> 1};
                if constexpr (kSum) __gtap_spawn_task_sum(lane, F_NQ, 0, child);
                else __gtap_spawn_task(lane, F_NQ, 0, kOffCounts + 8 * k, child);
                ++k;
            }
            if (k == 0) {  // dead end: no legal column in this row
                __gtap_return(lane, (int64_t)0);
                return;
            }
            if constexpr (kSum) {
                if (__gtap_set_state_for_join_sum(lane, 1, 0)) return;
            } else {
                lane.template frame_st<int32_t>(kOffK, k);
                if (__gtap_set_state_for_join(lane, 1, 0)) return;
            }
        }
            [[fallthrough]];
        case 1: {
            if constexpr (kSum) {
                __gtap_return(lane, __gtap_get_join_sum(lane));
            } else {
                const int32_t k = lane.template arg<int32_t>(kOffK);
                int64_t s = 0;
                for (int32_t i = 0; i < k; i += 2) {
                    int64_t a, b;
                    lane.template frame_ld2<int64_t>(kOffCounts + 8 * i, a, b);
                    s += a + (i + 1 < k ? b : 0);
                }
                __gtap_return(lane, s);
            }
            return;
        }
        default:
            GTAP_DEVICE_ERROR(*lane.dev, GTAP_ERROR_INVALID_TASKWAIT, lane.tid, 0, lane.state, 1);
            lane.finish();
            return;
    }
}

template <bool kSum>
template <class L>
__device__ __forceinline__ void ProgramT<kSum>::run(L& lane) {
    nq_task<kSum>(lane);
}

}  // namespace nq
}  // namespace gtap

namespace gtap {
namespace nq {

// The reference's evaluation program at its own settings (evaluation/2-comparison/nq/
// gtap_nq.cu:10-56): tasks down to row `cutoff`, a sequential search below it, detached
// spawns (no taskwait) and one shared int32 counter. The sequential search here is an
// explicit-stack depth-first walk; each task adds its count with one atomic (the reference
// adds 1 per solution). Unlike the reference, a task that reaches row N counts a solution
// even above the cutoff, so N <= cutoff + 1 is also correct.
struct CutoffArgs {
    int32_t n;
    int32_t cutoff;
    int32_t* answer;
};

__device__ __forceinline__ int32_t nq_serial(int32_t row, int32_t n, uint32_t cols, uint32_t left, uint32_t right) {
    const uint32_t mask = (n < 32) ? ((1u << n) - 1u) : 0xFFFFFFFFu;
    uint32_t avail[kMaxN + 1], cs[kMaxN + 1], ls[kMaxN + 1], rs[kMaxN + 1];
    int32_t count = 0;
    int32_t d = 0;
    cs[0] = cols;
    ls[0] = left;
    rs[0] = right;
    avail[0] = mask & ~(cols | left | right);
    if (row == n) return 1;
    while (d >= 0) {
        const uint32_t a = avail[d];
        if (a == 0u) {
            --d;
            continue;
        }
        const uint32_t p = a & (0u - a);
        avail[d] = a - p;
        if (row + d + 1 == n) {
            ++count;
            continue;
        }
        const uint32_t c = cs[d] | p, l = (ls[d] | p) << 1, r = (rs[d] | p) >> 1;
        ++d;
        cs[d] = c;
        ls[d] = l;
        rs[d] = r;
        avail[d] = mask & ~(c | l | r);
    }
    return count;
}

struct CutoffProgram {
    static constexpr int kQueues = 1;
    static constexpr bool kNoTaskwait = true;
    static constexpr int kWavesPerSimd = 3;
    static constexpr int kFrameWords = 4;  // row, cols, left, right
    static constexpr uint32_t kMaxFunc = F_NQ;
    static constexpr uint32_t kFrameBytes = 16;
    using Args = CutoffArgs;
    template <class L>
    __device__ __forceinline__ static void run(L& lane) {
        const CutoffArgs& g = *lane.args;
        const int32_t row = lane.template arg<int32_t>(0);
        const uint32_t cols = lane.template arg<uint32_t>(4);
        const uint32_t left = lane.template arg<uint32_t>(8);
        const uint32_t right = lane.template arg<uint32_t>(12);
        int32_t found = 0;
        if (row == g.n) {
            found = 1;
        } else if (row > g.cutoff) {
            found = nq_serial(row, g.n, cols, left, right);
        } else {
            const uint32_t mask = (g.n < 32) ? ((1u << g.n) - 1u) : 0xFFFFFFFFu;
            uint32_t avail = mask & ~(cols | left | right);
            while (avail) {
                const uint32_t p = avail & (0u - avail);
                avail -= p;
                __gtap_spawn_task_detached(lane, F_NQ, 0, NqArg{row + 1, cols | p, (left | p) << 1, (right | p) >> 1});
            }
        }
        if (found) atomic_add(g.answer, found);
        __gtap_finish_task(lane);
    }
};

}  // namespace nq
}  // namespace gtap

using NqProgram = gtap::nq::Program;
using NqSlotProgram = gtap::nq::SlotProgram;
static_assert(NqSlotProgram::kFrameBytes + gtap::kHdrBytes <= gtap::host::kBuiltinStride, "frame exceeds stride");

// Counts the n-queens placements of an n x n board as one persistent launch
// (reference program: evaluation/2-comparison/nq/gtap_nq.cu:57-99).
extern "C" int gtap_nqueens(int n, int64_t* result, gtap_run_stats* stats) {
    if (n < 1 || n > gtap::nq::kMaxN) return gtap::host::invalid_argument("gtap_nqueens: n must be in [1, 16]");
    gtap::nq::NqArg root{0, 0u, 0u, 0u};
    gtap_job job = gtap::host::make_job(gtap::nq::F_NQ, 0, &root, sizeof(root));
    int64_t r = 0;
    int st = gtap::host::run_single_job<NqProgram>(NqProgram::Args{n}, job, &r, stats, GTAP_TERM_JOBS);
    if (st == 0 && result) *result = r;
    return st;
}

// The same count with the result-slot join (each child stores its count into the parent's
// frame, then counts the join down), the form a slot-per-call lowering produces; work-
// stealing scheduler only.
extern "C" int gtap_nqueens_slots(int n, int64_t* result, gtap_run_stats* stats) {
    if (n < 1 || n > gtap::nq::kMaxN) return gtap::host::invalid_argument("gtap_nqueens_slots: n must be in [1, 16]");
    gtap::nq::NqArg root{0, 0u, 0u, 0u};
    gtap_job job = gtap::host::make_job(gtap::nq::F_NQ, 0, &root, sizeof(root));
    int64_t r = 0;
    int st = gtap::host::run_single_job<NqSlotProgram>(NqSlotProgram::Args{n}, job, &r, stats, GTAP_TERM_JOBS);
    if (st == 0 && result) *result = r;
    return st;
}

// The reference's n-queens benchmark at its own settings (evaluation/2-comparison/nq/
// gtap_nq.cu:57-99): tasks for rows 0..cutoff (the reference's default 7), a sequential search
// below, no taskwait, one int32 counter. Returns the solution count in *result.
extern "C" int gtap_nqueens_cutoff(int n, int cutoff, int64_t* result, gtap_run_stats* stats) {
    using namespace gtap;
    if (n < 1 || n > nq::kMaxN || cutoff < 0) return host::invalid_argument("gtap_nqueens_cutoff: n in [1, 16], cutoff >= 0");
    host::Runtime& R = host::rt();
    if (!R.initialized) {
        int st = gtap_initialize();
        if (st) return st;
    }
    // the counter: the launch's program control word, cleared by the per-launch reset
    int32_t* answer = R.ctl + CTL_PROGRAM * 32;
    nq::NqArg root{0, 0u, 0u, 0u};
    gtap_job job = host::make_job(nq::F_NQ, 0, &root, sizeof(root));
    int st = host::run_single_job<gtap::nq::CutoffProgram>(nq::CutoffArgs{n, cutoff, answer}, job, nullptr, stats,
                                                            GTAP_TERM_QUIESCE);
    if (st) return st;
    int32_t v = 0;
    GTAP_HIP_TRY(hipMemcpy(&v, answer, sizeof(v), hipMemcpyDeviceToHost));
    if (result) *result = v;
    return 0;
}
