// prog_tree.hip -- the synthetic binary-tree workload: every node of a complete binary tree
// is a task that spawns its two children, waits for them, then does its own memory and
// arithmetic work and stores one float64 result (thread-executed, one task per lane).
//
// Source program (the task shape of the reference's examples/binary_tree/binary_tree_thread.cu:63-120):
//
//   __device__ double node_work(int node, int mem_ops, int compute_iters) {
//       uint32_t seed = xorshift32(node ^ 0x9e3779b9);
//       double acc = 0;
//       for (int m = 0; m < mem_ops; ++m)                       // random reads of the input
//           acc += input[xorshift32(seed + m * 747796405) & (input_n - 1)];
//       uint32_t mix = (0x9e3779b9 ^ node) + (uint32_t)acc;
//       for (int it = 0; it < compute_iters; ++it) {            // arithmetic
//           double y = fma((double)xorshift32(seed + it * 747796405), 1.0000001192092896,
//                          0.9999999403953552);
//           mix ^= (uint32_t)bits(y);
//       }
//       return (double)mix;
//   }
//   #pragma gtap function
//   __device__ void tree_work(int node, int height, int mem_ops, int compute_iters) {
//       if (height == 0) { out[node] = node_work(node, ...); return; }
//       #pragma gtap task
//       tree_work(2 * node + 1, height - 1, ...);
//       #pragma gtap task
//       tree_work(2 * node + 2, height - 1, ...);
//       #pragma gtap taskwait
//       out[node] = node_work(node, ...);
//   }
//
// `(uint32_t)acc` converts a double that may be negative. In C++ that is undefined; the
// reference's GPU converts with saturation (negative and NaN -> 0, above 2^32-1 -> 2^32-1),
// and so does sat_u32 here, explicitly, so the result does not depend on the compiler.
#include <cstring>

#include "../device/gtap_abi.hpp"
#include "../host/gtap_launch.hpp"

namespace gtap {
namespace tree {

enum : uint32_t { F_NODE = 1 };
struct NodeArg {
    int32_t node;
    int32_t height;
    int32_t mem_ops;
    int32_t compute_iters;
};
struct TreeArgs {
    const double* input;
    uint32_t input_mask;  // input_n - 1 (input_n a power of two)
    double* out;
};

__host__ __device__ __forceinline__ uint32_t xs32(uint32_t x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

__device__ __forceinline__ uint32_t sat_u32(double v) {
    if (!(v > 0.0)) return 0u;  // negative, zero and NaN
    if (v >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

// Memory phase: 8 loads in flight at a time, summed in order.
__device__ __forceinline__ double node_work(const TreeArgs& g, int32_t node, int32_t mem_ops, int32_t compute_iters) {
    const uint32_t seed = xs32((uint32_t)node ^ 0x9e3779b9u);
    double acc = 0.0;
    constexpr int kBatch = 8;
    for (int32_t m0 = 0; m0 < mem_ops; m0 += kBatch) {
        double v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const uint32_t r = xs32(seed + (uint32_t)(m0 + j) * 747796405u);
            v[j] = (m0 + j < mem_ops) ? g.input[r & g.input_mask] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j)
            if (m0 + j < mem_ops) acc += v[j];
    }
    uint32_t mix = (0x9e3779b9u ^ (uint32_t)node) + sat_u32(acc);
    for (int32_t it = 0; it < compute_iters; ++it) {
        const double x = (double)xs32(seed + (uint32_t)it * 747796405u);
        const double y = __builtin_fma(x, 1.0000001192092896, 0.9999999403953552);
        mix ^= (uint32_t)__builtin_bit_cast(uint64_t, y);
    }
    return (double)mix;
}

constexpr int kUnusedSlot = 16;  // frame bytes 16..23

struct Program {
    static constexpr int kQueues = 1;
    static constexpr int kFrameWords = 4;  // node, height, mem_ops, compute_iters
    static constexpr uint32_t kMaxFunc = F_NODE;
    static constexpr uint32_t kFrameBytes = 24;
    using Args = TreeArgs;
    template <class L>
    __device__ __forceinline__ static void run(L& lane);
};

template <class L>
__device__ __forceinline__ void Program::run(L& lane) {
    const TreeArgs& g = *lane.args;
    const int32_t node = lane.template arg<int32_t>(0);
    const int32_t height = lane.template arg<int32_t>(4);
    const int32_t mem_ops = lane.template arg<int32_t>(8);
    const int32_t iters = lane.template arg<int32_t>(12);
    switch (__gtap_get_task_state(lane)) {
        case 0:
            if (height > 0) {
                // the children return nothing: their result slot is an unused frame word
                __gtap_spawn_task(lane, F_NODE, 0, kUnusedSlot, NodeArg{2 * node + 1, height - 1, mem_ops, iters});
                __gtap_spawn_task(lane, F_NODE, 0, kUnusedSlot, NodeArg{2 * node + 2, height - 1, mem_ops, iters});
                if (__gtap_set_state_for_join(lane, 1, 0)) return;
            }
            [[fallthrough]];
        case 1:
            g.out[node] = node_work(g, node, mem_ops, iters);
            __gtap_finish_task(lane);
            return;
        default:
            GTAP_DEVICE_ERROR(*lane.dev, GTAP_ERROR_INVALID_TASKWAIT, lane.tid, 0, lane.state, 1);
            __gtap_finish_task(lane);
            return;
    }
}

}  // namespace tree
}  // namespace gtap

using TreeProgram = gtap::tree::Program;
static_assert(TreeProgram::kFrameBytes + gtap::kHdrBytes <= gtap::host::kBuiltinStride, "frame exceeds stride");

// The binary-tree workload over a tree of the given height (2^(height+1) - 1 nodes). input
// (input_n float64, input_n a power of two) and out (one float64 per node) are device
// buffers. Replaces the program of examples/binary_tree/binary_tree_thread.cu:93-120.
extern "C" int gtap_tree(int32_t height, int32_t mem_ops, int32_t compute_iters, const double* input, int64_t input_n,
                         double* out, gtap_run_stats* stats) {
    using namespace gtap;
    if (height < 0 || height > 29 || mem_ops < 0 || compute_iters < 0 || !out || (mem_ops > 0 && !input) ||
        input_n < 1 || input_n > (int64_t)1 << 31 || (input_n & (input_n - 1)) != 0)
        return host::invalid_argument("gtap_tree: bad arguments (input_n must be a power of two)");
    tree::NodeArg root{0, height, mem_ops, compute_iters};
    gtap_job job = host::make_job(tree::F_NODE, 0, &root, sizeof(root));
    tree::TreeArgs a{input, (uint32_t)(input_n - 1), out};
    return host::run_single_job<TreeProgram>(a, job, nullptr, stats, GTAP_TERM_JOBS);
}

namespace {

uint64_t tree_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

// The tree's input: n float64 uniform in [-1, 1) from a splitmix64 stream (the reference
// draws them with std::mt19937_64, binary_tree_thread.cu:139-144).
extern "C" int gtap_tree_input(uint64_t seed, int64_t n, double* out) {
    if (n < 0 || (n > 0 && !out)) return gtap::host::invalid_argument("gtap_tree_input");
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t r = tree_mix(seed + 0x9E3779B97F4A7C15ULL * ((uint64_t)i + 1));
        out[i] = (double)(r >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    return 0;
}
