// prog_fib.hip -- fib as a thread-executed task program.
//
// Source program (the reference's evaluation/2-comparison/fib/gtap_fib.cu:10-22, with an
// int64 result):
//
//   #pragma gtap function
//   __device__ int64_t fib(int n) {
//       if (n < 2) return n;
//       int64_t a, b;
//       #pragma gtap task
//       a = fib(n - 1);
//       #pragma gtap task
//       b = fib(n - 2);
//       #pragma gtap taskwait
//       return a + b;
//   }
//
// Lowered form below, states 0 and 1, with the additive join: a and b are only summed after
// the taskwait, so each child adds its value to the parent's join word with the atomic that
// counts the join down, and the parent resumes with a + b. Frame = { int32 n; int32 pad; }.
#include "../device/gtap_abi.hpp"
#include "../host/gtap_launch.hpp"

namespace gtap {
namespace fib {

enum : uint32_t { F_FIB = 1 };
struct NArg {
    int32_t n;
    int32_t pad;
};

struct Program {
    static constexpr int kQueues = 1;
    static constexpr bool kSchedVariants = true;  // also built with the global-queue and Chase-Lev schedulers
    static constexpr int kWavesPerSimd = 4;
    static constexpr int kFrameWords = 2;  // n (+pad)
    static constexpr uint32_t kMaxFunc = F_FIB;
    static constexpr uint32_t kFrameBytes = 16;  // covers the 4 frame words loaded with a task
    struct Args {
        int32_t unused;
    };
    template <class L>
    __device__ __forceinline__ static void run(L& lane);
};

template <class L>
__device__ __forceinline__ void fib_task(L& lane) {
    switch (__gtap_get_task_state(lane)) {
        case 0: {
            const int32_t n = lane.template arg<int32_t>(0);
            if (n < 2) {
                __gtap_return(lane, (int64_t)n);
                return;
            }
            __gtap_spawn_task_sum(lane, F_FIB, 0, NArg{n - 1, 0});
            __gtap_spawn_task_sum(lane, F_FIB, 0, NArg{n - 2, 0});
            if (__gtap_set_state_for_join_sum(lane, 1, 0)) return;
        }
            [[fallthrough]];
        case 1:
            __gtap_return(lane, __gtap_get_join_sum(lane));
            return;
        default:
            GTAP_DEVICE_ERROR(*lane.dev, GTAP_ERROR_INVALID_TASKWAIT, lane.tid, 0, lane.state, 1);
            lane.finish();
            return;
    }
}

template <class L>
__device__ __forceinline__ void Program::run(L& lane) {
    fib_task(lane);
}

}  // namespace fib
}  // namespace gtap

using FibProgram = gtap::fib::Program;
static_assert(FibProgram::kFrameBytes + gtap::kHdrBytes <= gtap::host::kBuiltinStride, "frame exceeds stride");

// Runs fib(n) as one persistent launch (one job). Reference host flow:
// examples/fib/fib.cu:26-57.
extern "C" int gtap_fib(int n, int64_t* result, gtap_run_stats* stats) {
    if (n < 0 || n > 92) return gtap::host::invalid_argument("gtap_fib: n must be in [0, 92]");
    gtap_job job = gtap::host::make_job(gtap::fib::F_FIB, 0, &n, sizeof(n));
    int64_t r = 0;
    int st = gtap::host::run_single_job<FibProgram>(FibProgram::Args{0}, job, &r, stats, GTAP_TERM_JOBS);
    if (st == 0 && result) *result = r;
    return st;
}
