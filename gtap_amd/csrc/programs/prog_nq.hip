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
// Lowered frame: { int32 row; uint32 cols, diag, anti; int32 k; int32 pad[3]; int64 counts[16]; }.
// The first six words (the arguments and k) are the frame words loaded when the task starts,
// so the resumed task knows k without another load; counts[] is 16-byte aligned and read
// two at a time.
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

struct Program {
    static constexpr int kQueues = 1;
    static constexpr bool kSchedVariants = true;  // also built with the global-queue and Chase-Lev schedulers
    static constexpr int kFrameWords = 6;  // row, cols, diag, anti, k, pad
    static constexpr uint32_t kMaxFunc = F_NQ;
    static constexpr uint32_t kFrameBytes = kOffCounts + 8 * kMaxN;
    struct Args {
        int32_t n;
    };
    template <class L>
    __device__ __forceinline__ static void run(L& lane);
};

template <class L>
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
                __gtap_spawn_task(lane, F_NQ, 0, kOffCounts + 8 * k,
                                  NqArg{row + 1, cols | p, (diag | p) << 1, (anti | p) >> 1});
                ++k;
            }
            if (k == 0) {  // dead end: no legal column in this row
                __gtap_return(lane, (int64_t)0);
                return;
            }
            lane.template frame_st<int32_t>(kOffK, k);
            if (__gtap_set_state_for_join(lane, 1, 0)) return;
        }
            [[fallthrough]];
        case 1: {
            const int32_t k = lane.template arg<int32_t>(kOffK);
            int64_t s = 0;
            for (int32_t i = 0; i < k; i += 2) {
                int64_t a, b;
                lane.template frame_ld2<int64_t>(kOffCounts + 8 * i, a, b);
                s += a + (i + 1 < k ? b : 0);
            }
            __gtap_return(lane, s);
            return;
        }
        default:
            GTAP_DEVICE_ERROR(*lane.dev, GTAP_ERROR_INVALID_TASKWAIT, lane.tid, 0, lane.state, 1);
            lane.finish();
            return;
    }
}

template <class L>
__device__ __forceinline__ void Program::run(L& lane) {
    nq_task(lane);
}

}  // namespace nq
}  // namespace gtap

using NqProgram = gtap::nq::Program;
static_assert(NqProgram::kFrameBytes + gtap::kHdrBytes <= gtap::host::kBuiltinStride, "frame exceeds stride");

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
