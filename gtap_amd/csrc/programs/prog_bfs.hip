// prog_bfs.hip -- asynchronous (label-correcting) BFS as a thread-executed task program,
// with EPAQ routing of vertex tasks by out-degree bucket.
//
// Source program (shape of the reference's examples/asynchronous_bfs/asynchronous_bfs_thread.cu:19-55
// and evaluation/2-comparison/bfs/gtap_thread_bfs.cu:21-54):
//
//   #pragma gtap function
//   __device__ void bfs(int v, int s, int e, int d) {
//       if (depth[v] < d) return;                      // a shorter path already re-expanded v
//       if (e - s > SPLIT) {                           // split a long adjacency range into
//           int k = min(8, ceil((e - s) / SPLIT));     // k near-equal pieces
//           for (int i = 0; i < k; ++i) {
//               int ps = s + (e - s) * i / k, pe = s + (e - s) * (i + 1) / k;
//               #pragma gtap task queue(bucket(pe - ps))
//               bfs(v, ps, pe, d);
//           }
//           return;
//       }
//       for (int i = s; i < e; ++i) {
//           int u = col[i];
//           if (atomicMin(&depth[u], d + 1) > d + 1) {
//               #pragma gtap task queue(bucket(row[u + 1] - row[u]))
//               bfs(u, row[u], row[u + 1], d + 1);
//           }
//       }
//   }
//
// bucket(deg) = 0 for deg < 8, 1 for < 64, 2 for < 512, 3 otherwise (4 EPAQ queues); with one
// queue every task goes to queue 0. No task waits for its children (detached spawns); the
// launch ends by quiescence. Distances are the BFS hop counts whatever the schedule.
//
// With 4 queues, queue 3 (ranges of 512 edges and more) is a wave queue: its task runs on a
// whole wavefront, which reads the range 512 edges at a time (lane l takes edges b + l + 64 j,
// so each column load is one contiguous 256-byte read) and spawns every improved neighbour
// from the lane that found it. A range above kWaveSplitEdges (or an eighth of the deque
// capacity, if smaller) is first cut into up to 64 near-equal pieces, one spawned per lane.
// A hub's adjacency list is thus scanned by a few wave rounds instead of a tree of 64-edge
// lane tasks and their split tasks. One queue keeps the lane program for every range. (The
// reference cuts a long list into 32-edge tasks all spawned by one task,
// asynchronous_bfs_thread.cu:25-33; a bounded fan-out keeps one lane's spawn loop short.)
#include "../device/gtap_abi.hpp"
#include "../device/gtap_block.hpp"
#include "../host/gtap_launch.hpp"

namespace gtap {
namespace bfs {

enum : uint32_t { F_BFS = 1 };
struct BfsArg {
    int32_t v;
    int32_t s;
    int32_t e;
    int32_t d;
};
struct BfsArgs {
    const int32_t* row;
    const int32_t* col;
    int32_t* depth;
    int32_t nv;
    int32_t split;
    int64_t ne;  // column entries: a vertex's offsets must satisfy 0 <= row[u] <= row[u+1] <= ne
};

__host__ __device__ __forceinline__ uint32_t degree_bucket(int32_t deg) {
    return deg < 8 ? 0u : (deg < 64 ? 1u : (deg < 512 ? 2u : 3u));
}

constexpr int32_t kSplitFanout = 8;         // pieces per split of a lane task's range
constexpr int32_t kWaveSplitEdges = 2048;  // a wave task scans at most this many edges (and cap / 8)
constexpr int kWaveChunk = 8;               // edges per lane per step of a wave scan

template <int NQ>
struct Program {
    static constexpr int kQueues = NQ;
    static constexpr int kWaveQueue = NQ == 4 ? 3 : -1;
    static constexpr bool kNoTaskwait = true;  // detached spawns only
    static constexpr int kWavesPerSimd = 3;    // 3072 wavefront workers with 4 queues too (4 per SIMD measured slower)
    static constexpr int kFrameWords = 4;
    static constexpr uint32_t kMaxFunc = F_BFS;
    static constexpr uint32_t kFrameBytes = 16;
    using Args = BfsArgs;
    struct WaveShared {
        int32_t unused;
    };
    template <class L>
    __device__ __forceinline__ static void run(L& lane);
    template <class W>
    __device__ __forceinline__ static void run_wave(W& wt);
};

template <int NQ>
__device__ __forceinline__ uint32_t route(int32_t deg) {
    if constexpr (NQ == 1) {
        return 0u;
    } else {
        return degree_bucket(deg);
    }
}

template <int NQ, class L>
__device__ __forceinline__ void bfs_task(L& lane) {
    const BfsArgs& g = *lane.args;
    const int32_t v = lane.template arg<int32_t>(0);
    const int32_t s = lane.template arg<int32_t>(4);
    const int32_t e = lane.template arg<int32_t>(8);
    const int32_t d = lane.template arg<int32_t>(12);
    if (ld32(reinterpret_cast<const int32_t*>(g.depth) + v) < d) {  // stale: v was reached by a shorter path meanwhile
        __gtap_finish_task(lane);
        return;
    }
    if (e - s > g.split) {
        // up to kSplitFanout near-equal pieces, each routed by its own length
        const int64_t len = (int64_t)e - s;
        int32_t k = (int32_t)((len + g.split - 1) / g.split);
        if (k > kSplitFanout) k = kSplitFanout;
        for (int32_t i = 0; i < k; ++i) {
            const int32_t ps = s + (int32_t)(len * i / k);
            const int32_t pe = s + (int32_t)(len * (i + 1) / k);
            __gtap_spawn_task_detached(lane, F_BFS, route<NQ>(pe - ps), BfsArg{v, ps, pe, d});
        }
        __gtap_finish_task(lane);
        return;
    }
    // Edges in chunks of kChunkEdges: the column loads, the depth probes and the atomicMins of
    // a chunk are issued back to back, so their latencies overlap instead of adding up edge
    // by edge. A plain load of depth[u] filters the atomic: depths only decrease, so a stale
    // cached value is never below the true one and can only cost an unneeded atomic. A chunk
    // is 16 edges: fewer dependent chains per task than 8 (1.33 -> 1.29 ms with one queue,
    // 1.43 -> 1.40 ms with four).
    constexpr int kChunkEdges = 16;
    const int32_t nd = d + 1;
    for (int32_t b = s; b < e; b += kChunkEdges) {
        int32_t u[kChunkEdges];
#pragma unroll
        for (int j = 0; j < kChunkEdges; ++j) u[j] = (b + j < e) ? g.col[b + j] : -1;
        bool bad = false;
#pragma unroll
        for (int j = 0; j < kChunkEdges; ++j) {
            if (u[j] >= g.nv || u[j] < -1) {
                bad = true;
                u[j] = -1;
            }
        }
        if (bad) {
            GTAP_DEVICE_ERROR(*lane.dev, GTAP_ERROR_INVALID_ARGUMENT, lane.tid, -1, -1, g.nv);
            break;
        }
        int32_t old[kChunkEdges];
#pragma unroll
        for (int j = 0; j < kChunkEdges; ++j) old[j] = (u[j] >= 0) ? g.depth[u[j]] : 0;
#pragma unroll
        for (int j = 0; j < kChunkEdges; ++j) {
            if (u[j] >= 0 && old[j] > nd)
                old[j] = __hip_atomic_fetch_min(g.depth + u[j], nd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else
                old[j] = 0;
        }
        // one spawn site for the improved vertices of the chunk (keeps the register count of the
        // spawn path single)
        uint32_t hits = 0;
#pragma unroll
        for (int j = 0; j < kChunkEdges; ++j) hits |= (old[j] > nd ? 1u : 0u) << j;
        while (hits) {
            const int j = __builtin_ctz(hits);
            hits &= hits - 1u;
            int32_t uj = u[0];
#pragma unroll
            for (int k = 1; k < kChunkEdges; ++k) uj = (j == k) ? u[k] : uj;
            const int32_t us = g.row[uj];
            const int32_t ue = g.row[uj + 1];
            if (us < 0 || ue < us || (int64_t)ue > g.ne) {  // malformed offsets: never read past col
                GTAP_DEVICE_ERROR(*lane.dev, GTAP_ERROR_INVALID_ARGUMENT, lane.tid, -1, uj, g.nv);
                continue;
            }
            __gtap_spawn_task_detached(lane, F_BFS, route<NQ>(ue - us), BfsArg{uj, us, ue, nd});
        }
    }
    __gtap_finish_task(lane);
}

template <int NQ>
template <class L>
__device__ __forceinline__ void Program<NQ>::run(L& lane) {
    bfs_task<NQ>(lane);
}

// The wave form of the same task (queue 3 with 4 queues): all 64 lanes enter with one (v, s, e, d).
template <int NQ>
template <class W>
__device__ __forceinline__ void Program<NQ>::run_wave(W& wt) {
    const BfsArgs& g = *wt.args;
    const int32_t v = wt.template arg<int32_t>(0);
    const int32_t s = wt.template arg<int32_t>(4);
    const int32_t e = wt.template arg<int32_t>(8);
    const int32_t d = wt.template arg<int32_t>(12);
    if (wt.func != F_BFS) {
        if (wt.lane == 0) GTAP_DEVICE_ERROR(*wt.dev, GTAP_ERROR_INVALID_FUNCTION, wt.tid, wt.queue, wt.func, F_BFS);
        wt.finish();
        return;
    }
    if ((int32_t)uniform((uint32_t)ld32(reinterpret_cast<const int32_t*>(g.depth) + v)) < d) {  // stale
        wt.finish();
        return;
    }
    // a scan spawns at most one task per edge in one round: an eighth of the deque capacity
    // stays within what a round may append (the deque keeps up to half its capacity published
    // and a margin free)
    const int32_t wsplit = kWaveSplitEdges < (int32_t)(wt.dev->cap / 8) ? kWaveSplitEdges : (int32_t)(wt.dev->cap / 8);
    if (e - s > wsplit) {
        // up to 64 near-equal pieces, lane i spawning piece i
        const int64_t len = (int64_t)e - s;
        int32_t k = (int32_t)((len + wsplit - 1) / wsplit);
        if (k > kWave) k = kWave;
        if (wt.lane < k) {
            const int32_t ps = s + (int32_t)(len * wt.lane / k);
            const int32_t pe = s + (int32_t)(len * (wt.lane + 1) / k);
            wt.spawn_detached_lane(F_BFS, route<NQ>(pe - ps), BfsArg{v, ps, pe, d});
        }
        wt.finish();
        return;
    }
    const int32_t nd = d + 1;
    const int32_t l = wt.lane;
    for (int32_t b = s; b < e; b += kWave * kWaveChunk) {
        int32_t u[kWaveChunk];
#pragma unroll
        for (int j = 0; j < kWaveChunk; ++j) {
            const int32_t i = b + j * kWave + l;
            u[j] = (i < e) ? g.col[i] : -1;
        }
        bool bad = false;
#pragma unroll
        for (int j = 0; j < kWaveChunk; ++j) {
            if (u[j] >= g.nv || u[j] < -1) {
                bad = true;
                u[j] = -1;
            }
        }
        if (ballot(bad)) {
            if (bad) GTAP_DEVICE_ERROR(*wt.dev, GTAP_ERROR_INVALID_ARGUMENT, wt.tid, -1, -1, g.nv);
            break;
        }
        int32_t old[kWaveChunk];
#pragma unroll
        for (int j = 0; j < kWaveChunk; ++j) old[j] = (u[j] >= 0) ? g.depth[u[j]] : 0;
#pragma unroll
        for (int j = 0; j < kWaveChunk; ++j) {
            if (u[j] >= 0 && old[j] > nd)
                old[j] = __hip_atomic_fetch_min(g.depth + u[j], nd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else
                old[j] = 0;
        }
        uint32_t hits = 0;
#pragma unroll
        for (int j = 0; j < kWaveChunk; ++j) hits |= (old[j] > nd ? 1u : 0u) << j;
        while (hits) {
            const int j = __builtin_ctz(hits);
            hits &= hits - 1u;
            int32_t uj = u[0];
#pragma unroll
            for (int k = 1; k < kWaveChunk; ++k) uj = (j == k) ? u[k] : uj;
            const int32_t us = g.row[uj];
            const int32_t ue = g.row[uj + 1];
            if (us < 0 || ue < us || (int64_t)ue > g.ne) {
                GTAP_DEVICE_ERROR(*wt.dev, GTAP_ERROR_INVALID_ARGUMENT, wt.tid, -1, uj, g.nv);
                continue;
            }
            wt.spawn_detached_lane(F_BFS, route<NQ>(ue - us), BfsArg{uj, us, ue, nd});
        }
    }
    wt.finish();
}

}  // namespace bfs
}  // namespace gtap

namespace gtap {
namespace bfs {

// Block-mode form (examples/asynchronous_bfs/asynchronous_bfs_block.cu:17-32): one vertex task
// runs on a whole 256-thread workgroup, thread t taking edge s + t of the vertex's range; each
// thread that lowers a neighbour's depth spawns that neighbour's task. A range longer than
// kBlockEdges is split in halves first (a workgroup invocation spawns at most kBlockMaxSpawn
// tasks; the reference walks the whole list with a stride instead).
constexpr int32_t kBlockEdges = 128;

struct BlockProgram {
    static constexpr int kFrameWords = 4;
    static constexpr uint32_t kMaxFunc = F_BFS;
    static constexpr uint32_t kFrameBytes = 16;
    using Args = BfsArgs;
    struct Shared {
        int32_t unused;
    };
    template <class B>
    __device__ __forceinline__ static void run(B& blk) {
        const BfsArgs& g = *blk.args;
        const int32_t v = blk.template arg<int32_t>(0);
        const int32_t s = blk.template arg<int32_t>(4);
        const int32_t e = blk.template arg<int32_t>(8);
        const int32_t d = blk.template arg<int32_t>(12);
        const uint32_t t = threadIdx.x;
        const bool stale = ld32(reinterpret_cast<const int32_t*>(g.depth) + v) < d;
        if (!stale) {
            if (e - s > kBlockEdges) {
                if (t == 0) {
                    const int32_t m = s + (e - s) / 2;
                    blk.spawn_detached(F_BFS, BfsArg{v, s, m, d});
                    blk.spawn_detached(F_BFS, BfsArg{v, m, e, d});
                }
            } else if ((int32_t)t < e - s) {
                const int32_t u = g.col[s + (int32_t)t];
                if (u < 0 || u >= g.nv) {
                    GTAP_DEVICE_ERROR(*blk.dev, GTAP_ERROR_INVALID_ARGUMENT, blk.tid, -1, u, g.nv);
                } else if (g.depth[u] > d + 1 &&
                           __hip_atomic_fetch_min(g.depth + u, d + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > d + 1) {
                    const int32_t us = g.row[u], ue = g.row[u + 1];
                    if (us < 0 || ue < us || (int64_t)ue > g.ne)
                        GTAP_DEVICE_ERROR(*blk.dev, GTAP_ERROR_INVALID_ARGUMENT, blk.tid, -1, u, g.nv);
                    else
                        blk.spawn_detached(F_BFS, BfsArg{u, us, ue, d + 1});
                }
            }
        }
        blk.finish();
    }
};
static_assert(kBlockEdges <= (int32_t)kBlockMaxSpawn, "one invocation spawns at most kBlockMaxSpawn tasks");

}  // namespace bfs
}  // namespace gtap

using Bfs1 = gtap::bfs::Program<1>;
using Bfs4 = gtap::bfs::Program<4>;

namespace {

__global__ void bfs_init_depth(int32_t* depth, int32_t nv, int32_t source) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = i; v < nv; v += stride) depth[v] = (v == source) ? 0 : 0x3f3f3f3f;
}

}  // namespace

// BFS from `source` over a CSR graph already in device memory (row: nv+1 int32 offsets,
// col: ne int32 columns). depth (nv int32, device) receives the hop counts, 0x3f3f3f3f for
// unreached vertices. num_queues = 1 (no EPAQ) or 4 (EPAQ by out-degree bucket).
// Replaces the program of examples/asynchronous_bfs/asynchronous_bfs_thread.cu:51-141.
extern "C" int gtap_bfs(const int32_t* row, const int32_t* col, int32_t nv, int64_t ne, int32_t source,
                        int32_t* depth, int32_t num_queues, int32_t split, gtap_run_stats* stats) {
    using namespace gtap;
    if (!row || !col || !depth || nv <= 0 || source < 0 || source >= nv || ne < 0)
        return host::invalid_argument("gtap_bfs: bad graph arguments");
    if (num_queues != 1 && num_queues != 4) return host::invalid_argument("gtap_bfs: num_queues must be 1 or 4");
    if (split <= 0) split = 64;
    host::Runtime& R = host::rt();
    if (!R.initialized) {
        int st = gtap_initialize();
        if (st) return st;
    }
    hipLaunchKernelGGL(bfs_init_depth, dim3(1024), dim3(256), 0, R.stream, depth, nv, source);
    GTAP_HIP_TRY(hipGetLastError());
    int32_t rs[2];
    GTAP_HIP_TRY(hipMemcpy(rs, row + source, 2 * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (rs[0] < 0 || rs[1] < rs[0] || (int64_t)rs[1] > ne) return host::invalid_argument("gtap_bfs: bad source row offsets");
    bfs::BfsArg root{source, rs[0], rs[1], 0};
    const uint32_t q = (num_queues == 4) ? bfs::degree_bucket(rs[1] - rs[0]) : 0u;
    gtap_job job = host::make_job(bfs::F_BFS, q, &root, sizeof(root));
    bfs::BfsArgs a{row, col, depth, nv, split, ne};
    if (num_queues == 4) return host::run_single_job<Bfs4>(a, job, nullptr, stats, GTAP_TERM_QUIESCE);
    return host::run_single_job<Bfs1>(a, job, nullptr, stats, GTAP_TERM_QUIESCE);
}

// The block-mode BFS (one vertex task per 256-thread workgroup) over a CSR graph in device
// memory; same arguments and result as gtap_bfs. Replaces the program of
// examples/asynchronous_bfs/asynchronous_bfs_block.cu:17-37.
extern "C" int gtap_bfs_block(const int32_t* row, const int32_t* col, int32_t nv, int64_t ne, int32_t source,
                              int32_t* depth, gtap_run_stats* stats) {
    using namespace gtap;
    if (!row || !col || !depth || nv <= 0 || source < 0 || source >= nv || ne < 0)
        return host::invalid_argument("gtap_bfs_block: bad graph arguments");
    host::Runtime& R = host::rt();
    if (!R.initialized) {
        int st = gtap_initialize();
        if (st) return st;
    }
    hipLaunchKernelGGL(bfs_init_depth, dim3(1024), dim3(256), 0, R.stream, depth, nv, source);
    GTAP_HIP_TRY(hipGetLastError());
    int32_t rs[2];
    GTAP_HIP_TRY(hipMemcpy(rs, row + source, 2 * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (rs[0] < 0 || rs[1] < rs[0] || (int64_t)rs[1] > ne)
        return host::invalid_argument("gtap_bfs_block: bad source row offsets");
    bfs::BfsArg root{source, rs[0], rs[1], 0};
    gtap_job job = host::make_job(bfs::F_BFS, 0, &root, sizeof(root));
    bfs::BfsArgs a{row, col, depth, nv, 0, ne};
    return host::run_single_block_job<bfs::BlockProgram>(a, job, nullptr, stats, GTAP_TERM_QUIESCE);
}

// Host-pointer convenience: copies the CSR in, runs gtap_bfs, copies depth out.
extern "C" int gtap_bfs_host(const int32_t* row, const int32_t* col, int32_t nv, int64_t ne, int32_t source,
                             int32_t* depth_out, int32_t num_queues, int32_t split, gtap_run_stats* stats) {
    if (!row || !col || !depth_out || nv <= 0 || ne < 0) return gtap::host::invalid_argument("gtap_bfs_host");
    if (int st = gtap::host::check_csr_rows(row, nv, ne, "gtap_bfs_host")) return st;
    int32_t *drow = nullptr, *dcol = nullptr, *ddep = nullptr;
    GTAP_HIP_TRY(hipMalloc(&drow, (size_t)(nv + 1) * sizeof(int32_t)));
    GTAP_HIP_TRY(hipMalloc(&dcol, (size_t)(ne > 0 ? ne : 1) * sizeof(int32_t)));
    GTAP_HIP_TRY(hipMalloc(&ddep, (size_t)nv * sizeof(int32_t)));
    int st = 0;
    hipError_t e = hipMemcpy(drow, row, (size_t)(nv + 1) * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && ne > 0) e = hipMemcpy(dcol, col, (size_t)ne * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) st = gtap::host::hip_status(e, "gtap_bfs_host copy-in");
    if (!st) st = gtap_bfs(drow, dcol, nv, ne, source, ddep, num_queues, split, stats);
    if (!st) {
        e = hipMemcpy(depth_out, ddep, (size_t)nv * sizeof(int32_t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = gtap::host::hip_status(e, "gtap_bfs_host copy-out");
    }
    (void)hipFree(drow);
    (void)hipFree(dcol);
    (void)hipFree(ddep);
    return st;
}
