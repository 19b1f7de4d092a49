// prog_stream.hip -- the mixed job stream: fib, n-queens and mergesort jobs in one resident
// persistent kernel (batched: all jobs in one launch; served: jobs streamed in by a host
// loop while the kernel runs), and the plain form, one launch per job.
//
// Task program (two EPAQ queues: queue 0 = per-lane tasks, queue 1 = wave-cooperative tasks):
//   fib(n)                       lane  (evaluation/2-comparison/fib/gtap_fib.cu:10-22, int64,
//                                       children summed by the additive join)
//   nq(row, cols, diag, anti, N) lane  (evaluation/2-comparison/nq/gtap_nq.cu:25-55 shape,
//                                       taskwait + int64 sum by the additive join, N per job)
//   sort_root(arr, tmp, n)       lane  spawns msort on queue 1, taskwait; spawns csum over
//                                      chunks of the sorted keys, taskwait; returns their sum
//   msort(a, t, n, into_t)       wave  2-way mergesort of examples/mergesort/mergesort.cu:34-50:
//                                      n <= 512 -> the wave sorts in LDS; else two msort
//                                      children (ping-pong buffers), taskwait, then merge
//   merge(a, na, b, nb, dst)     wave  cilkmerge split of gtap_cilksort.cu:89-127 down to
//                                      4096 keys, merged by the wave with merge-path tiles
//   csum(keys, n, pos0, acc)     wave  adds sum_i mix(pos0 + i, keys[i]) to the sort root's
//                                      64-bit accumulator with one atomic
//
// A sort job's result is the checksum of its sorted keys, sum over positions i of
// mix((i << 32) ^ key[i]) mod 2^64 with mix the SplitMix64 finalizer (Steele, Lea, Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014): every key is checked at
// its position, and the sum is computed in parallel.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../device/gtap_abi.hpp"
#include "../host/gtap_launch.hpp"

namespace gtap {
namespace stream {

enum : uint32_t { F_FIB = 1, F_NQ = 2, F_SROOT = 3, F_MSORT = 4, F_MERGE = 5, F_CSUM = 6 };
enum : uint32_t { Q_LANE = 0, Q_WAVE = 1 };
constexpr int kWaveLeaf = 512;        // keys sorted by one wave in LDS
constexpr int kWaveMergeLeaf = 4096;  // default: keys merged by one wave (the tile loop takes up to 16384)
constexpr int kMaxWaveMergeLeaf = 16384;

struct FibArg {
    int32_t n;
    int32_t pad;
};
struct NqArg {
    int32_t row;
    uint32_t cols;
    uint32_t diag;
    uint32_t anti;
    int32_t n;
    int32_t pad;
};
struct SortRootArg {
    uint32_t* arr;
    uint32_t* tmp;
    int64_t n;
};
struct MsortArg {
    uint32_t* a;
    uint32_t* t;
    int64_t n;
    int32_t into_t;
    int32_t pad;
};
struct MergeArg {
    const uint32_t* a;
    const uint32_t* b;
    uint32_t* dst;
    int64_t na;
    int64_t nb;
};
struct CsumArg {
    const uint32_t* keys;
    int64_t n;
    int64_t pos0;
    uint64_t* acc;
};
constexpr int kDummySlot = 40;  // result slot of void children (never read)
constexpr int kAccOff = 24;     // the sort root's checksum accumulator (after its arguments)
constexpr int64_t kCsumChunk = 16384;  // keys per checksum task

// The checksum term of key `k` at position `pos`.
__host__ __device__ __forceinline__ uint64_t sort_mix(uint64_t pos, uint32_t k) {
    uint64_t z = (pos << 32) ^ (uint64_t)k;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct StreamArgs {
    int32_t merge_leaf;  // merges of at most this many keys run on one wave (512..16384)
};

struct Program {
    static constexpr int kQueues = 2;
    static constexpr int kWaveQueue = Q_WAVE;
    // 3 waves per SIMD (3072 workers): the sort leaf of 512 keys and a 64-entry staging area
    // keep a workgroup's LDS at 50 KB (3 per CU); at 2 waves per SIMD the batched stream took
    // 81.6 ms against 65.2 ms
    static constexpr int kWavesPerSimd = 3;
    static constexpr int kStageEntries = 64;
    static constexpr int kFrameWords = 10;
    static constexpr uint32_t kMaxFunc = F_CSUM;
    static constexpr uint32_t kFrameBytes = 48;  // the largest arguments (merge: 40 bytes) and the dummy slot
    using Args = StreamArgs;
    struct WaveShared {
        uint32_t keys[kWaveLeaf];
        int32_t split[kMaxWaveMergeLeaf / kWaveLeaf + 2];
    };
    template <class L>
    __device__ __forceinline__ static void run(L& lane);
    template <class W>
    __device__ __forceinline__ static void run_wave(W& wt);
};

// ---------------------------------------------------------------- lane tasks

template <class L>
__device__ __forceinline__ void fib_task(L& lane) {
    switch (lane.get_state()) {
        case 0: {
            const int32_t n = lane.template arg<int32_t>(0);
            if (n < 2) {
                lane.ret((int64_t)n);
                return;
            }
            lane.spawn_sum(F_FIB, Q_LANE, FibArg{n - 1, 0});
            lane.spawn_sum(F_FIB, Q_LANE, FibArg{n - 2, 0});
            if (lane.taskwait_sum(1, Q_LANE)) return;
        }
            [[fallthrough]];
        case 1:
            lane.ret(lane.join_sum());
            return;
        default:
            lane.finish();
            return;
    }
}

template <class L>
__device__ __forceinline__ void nq_task(L& lane) {
    switch (lane.get_state()) {
        case 0: {
            const int32_t row = lane.template arg<int32_t>(0);
            const uint32_t cols = lane.template arg<uint32_t>(4);
            const uint32_t diag = lane.template arg<uint32_t>(8);
            const uint32_t anti = lane.template arg<uint32_t>(12);
            const int32_t N = lane.template arg<int32_t>(16);
            if (row == N) {
                lane.ret((int64_t)1);
                return;
            }
            const uint32_t mask = (N < 32) ? ((1u << N) - 1u) : 0xFFFFFFFFu;
            uint32_t avail = mask & ~(cols | diag | anti);
            int32_t k = 0;
            while (avail) {
                const uint32_t p = avail & (0u - avail);
                avail -= p;
                lane.spawn_sum(F_NQ, Q_LANE, NqArg{row + 1, cols | p, (diag | p) << 1, (anti | p) >> 1, N, 0});
                ++k;
            }
            if (k == 0) {
                lane.ret((int64_t)0);
                return;
            }
            if (lane.taskwait_sum(1, Q_LANE)) return;
        }
            [[fallthrough]];
        case 1:
            lane.ret(lane.join_sum());
            return;
        default:
            lane.finish();
            return;
    }
}

template <class L>
__device__ __forceinline__ void sort_root_task(L& lane) {
    uint32_t* arr = lane.template arg<uint32_t*>(0);
    uint32_t* tmp = lane.template arg<uint32_t*>(8);
    const int64_t n = lane.template arg<int64_t>(16);
    switch (lane.get_state()) {
        case 0:
            if (n <= 1) {  // nothing to sort; the checksum of at most one key
                lane.ret(n == 1 ? (int64_t)sort_mix(0, ld32(arr)) : (int64_t)0);
                return;
            }
            lane.spawn(F_MSORT, Q_WAVE, kDummySlot, MsortArg{arr, tmp, n, 0, 0});
            if (lane.taskwait(1, Q_LANE)) return;
            [[fallthrough]];
        case 1: {
            // the sorted keys' checksum: wave tasks over chunks add into this record's accumulator
            lane.template frame_st<uint64_t>(kAccOff, 0ull);
            uint64_t* acc = reinterpret_cast<uint64_t*>(lane.rec + kHdrBytes + kAccOff);
            for (int64_t p = 0; p < n; p += kCsumChunk)
                lane.spawn(F_CSUM, Q_WAVE, kDummySlot, CsumArg{arr + p, n - p < kCsumChunk ? n - p : kCsumChunk, p, acc});
            if (lane.taskwait(2, Q_LANE)) return;
        }
            [[fallthrough]];
        default:
            lane.ret((int64_t)lane.template frame_ld<uint64_t>(kAccOff));
            return;
    }
}

template <class L>
__device__ __forceinline__ void Program::run(L& lane) {
    switch (lane.func) {
        case F_FIB: fib_task(lane); return;
        case F_NQ: nq_task(lane); return;
        case F_SROOT: sort_root_task(lane); return;
        default:
            // wave functions never arrive in a lane batch
            GTAP_DEVICE_ERROR(*lane.dev, GTAP_ERROR_INVALID_FUNCTION, lane.tid, lane.queue, lane.func, F_SROOT);
            lane.finish();
            return;
    }
}

// ---------------------------------------------------------------- wave tasks

// keys live in device memory (the caller's arrays and the scratch buffer)
__device__ __forceinline__ uint32_t ldk(const uint32_t* p) { return gld32(p); }
__device__ __forceinline__ void stk(uint32_t* p, uint32_t v) { gst32(p, v); }

// Sort src[0..n) (n <= kWaveLeaf) into dst with a bitonic network over kWaveLeaf LDS slots.
template <class W>
__device__ void wave_leaf_sort(W& wt, const uint32_t* src, uint32_t* dst, int64_t n) {
    uint32_t* k = wt.sh->keys;
    const uint32_t l = (uint32_t)wt.lane;
    for (uint32_t i = l; i < (uint32_t)kWaveLeaf; i += kWave) k[i] = (int64_t)i < n ? ldk(src + i) : 0xFFFFFFFFu;
    wave_sync();
    for (uint32_t size = 2; size <= (uint32_t)kWaveLeaf; size <<= 1) {
        for (uint32_t j = size >> 1; j > 0; j >>= 1) {
            for (uint32_t p = l; p < (uint32_t)kWaveLeaf / 2; p += kWave) {
                const uint32_t lo = 2 * j * (p / j) + (p % j);
                const uint32_t hi = lo + j;
                const uint32_t x = k[lo], y = k[hi];
                const bool up = (lo & size) == 0;
                if ((x > y) == up) {
                    k[lo] = y;
                    k[hi] = x;
                }
            }
            wave_sync();
        }
    }
    for (uint32_t i = l; (int64_t)i < n; i += kWave) stk(dst + i, k[i]);
    wave_sync();
}

template <class W>
__device__ void wave_copy(W& wt, const uint32_t* src, uint32_t* dst, int64_t n) {
    for (int64_t i = wt.lane; i < n; i += kWave) stk(dst + i, ldk(src + i));
}

// Number of `a` elements among the first d outputs (a first on ties), 32-bit indices.
template <class LDA, class LDB>
__device__ __forceinline__ int32_t merge_path32(LDA lda, int32_t na, LDB ldb, int32_t nb, int32_t d) {
    int32_t lo = d > nb ? d - nb : 0;
    int32_t hi = d < na ? d : na;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (lda(mid) <= ldb(d - 1 - mid)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Merge a[0..na) and b[0..nb) into dst (na + nb <= kMaxWaveMergeLeaf) by tiles of kWaveLeaf
// outputs, kWaveLeaf / 64 per lane, staged through the wave's LDS tile.
template <class W>
__device__ void wave_leaf_merge(W& wt, const uint32_t* a, int32_t na, const uint32_t* b, int32_t nb, uint32_t* dst) {
    typename Program::WaveShared* sh = wt.sh;
    const int32_t l = wt.lane;
    const int32_t total = na + nb;
    auto ga = [a](int32_t i) { return ldk(a + i); };
    auto gb = [b](int32_t i) { return ldk(b + i); };
    constexpr int32_t kTileOut = kWaveLeaf;
    constexpr int32_t kPer = kTileOut / kWave;
    const int32_t ntiles = (total + kTileOut - 1) / kTileOut;  // <= kMaxWaveMergeLeaf / kTileOut
    static_assert(kMaxWaveMergeLeaf / kTileOut + 1 <= (int)(sizeof(sh->split) / sizeof(sh->split[0])), "split table");
    if (l <= ntiles) {
        const int32_t d = (l * kTileOut < total) ? l * kTileOut : total;
        sh->split[l] = merge_path32(ga, na, gb, nb, d);
    }
    wave_sync();
    for (int32_t ti = 0; ti < ntiles; ++ti) {
        const int32_t d0 = ti * kTileOut;
        const int32_t d1 = (d0 + kTileOut < total) ? d0 + kTileOut : total;
        const int32_t i0 = (int32_t)sh->split[ti], i1 = (int32_t)sh->split[ti + 1];
        const int32_t la = i1 - i0, lb = (d1 - i1) - (d0 - i0);
        const int32_t j0 = d0 - i0;
        wave_sync();
        for (int32_t i = l; i < la; i += kWave) sh->keys[i] = ldk(a + i0 + i);
        for (int32_t i = l; i < lb; i += kWave) sh->keys[la + i] = ldk(b + j0 + i);
        wave_sync();
        const uint32_t* A = sh->keys;
        const uint32_t* Bv = sh->keys + la;
        const int32_t len = d1 - d0;
        const int32_t od = l * kPer;
        uint32_t outv[kPer];
        int32_t cnt = 0;
        if (od < len) {
            auto la_ = [A](int32_t i) { return A[i]; };
            auto lb_ = [Bv](int32_t i) { return Bv[i]; };
            int32_t i = merge_path32(la_, la, lb_, lb, od);
            int32_t j = od - i;
            cnt = (od + kPer < len) ? kPer : len - od;
#pragma unroll
            for (int32_t o = 0; o < kPer; ++o) {
                if (o < cnt) {
                    const bool takeA = i < la && (j >= lb || A[i] <= Bv[j]);
                    outv[o] = takeA ? A[i] : Bv[j];
                    i += takeA ? 1 : 0;
                    j += takeA ? 0 : 1;
                }
            }
        }
        wave_sync();  // every lane has finished reading the inputs
#pragma unroll
        for (int32_t o = 0; o < kPer; ++o)
            if (o < cnt) sh->keys[od + o] = outv[o];
        wave_sync();
        for (int32_t o = l; o < len; o += kWave) stk(dst + d0 + o, sh->keys[o]);
    }
    wave_sync();
}

// gtap_cilksort.cu:40-50
__device__ __forceinline__ int64_t binary_search_gm(const uint32_t* x, int64_t len, uint32_t val) {
    int64_t low = -1, high = len;
    while (low + 1 < high) {
        const int64_t mid = (low + high) / 2;
        if (ldk(x + mid) <= val) low = mid;
        else high = mid;
    }
    return low;
}

template <class W>
__device__ void msort_task(W& wt) {
    uint32_t* a = wt.template arg<uint32_t*>(0);
    uint32_t* t = wt.template arg<uint32_t*>(8);
    const int64_t n = wt.template arg<int64_t>(16);
    const int32_t into_t = wt.template arg<int32_t>(24);
    uint32_t* dst = into_t ? t : a;
    const int64_t h = n / 2;
    switch (wt.get_state()) {
        case 0:
            if (n <= kWaveLeaf) {
                wave_leaf_sort(wt, a, dst, n);
                wt.finish();
                return;
            }
            wt.spawn(F_MSORT, Q_WAVE, kDummySlot, MsortArg{a, t, h, !into_t, 0});
            wt.spawn(F_MSORT, Q_WAVE, kDummySlot, MsortArg{a + h, t + h, n - h, !into_t, 0});
            if (wt.taskwait(1, Q_WAVE)) return;
            [[fallthrough]];
        case 1: {
            const uint32_t* src = into_t ? a : t;  // where the sorted halves are
            if (n <= wt.args->merge_leaf) {
                wave_leaf_merge(wt, src, (int32_t)h, src + h, (int32_t)(n - h), dst);
                wt.finish();
                return;
            }
            wt.spawn(F_MERGE, Q_WAVE, kDummySlot, MergeArg{src, src + h, dst, h, n - h});
            if (wt.taskwait(2, Q_WAVE)) return;
        }
            [[fallthrough]];
        default:
            wt.finish();
            return;
    }
}

template <class W>
__device__ void merge_task(W& wt) {
    const uint32_t* a = wt.template arg<const uint32_t*>(0);
    const uint32_t* b = wt.template arg<const uint32_t*>(8);
    uint32_t* dst = wt.template arg<uint32_t*>(16);
    int64_t na = wt.template arg<int64_t>(24);
    int64_t nb = wt.template arg<int64_t>(32);
    if (wt.get_state() != 0) {
        wt.finish();
        return;
    }
    if (na < nb) {
        const uint32_t* tp = a;
        a = b;
        b = tp;
        const int64_t tn = na;
        na = nb;
        nb = tn;
    }
    if (nb == 0) {
        wave_copy(wt, a, dst, na);
        wt.finish();
        return;
    }
    if (na + nb <= wt.args->merge_leaf) {
        wave_leaf_merge(wt, a, (int32_t)na, b, (int32_t)nb, dst);
        wt.finish();
        return;
    }
    const int64_t a_split = (na + 1) / 2;
    const int64_t b_split = binary_search_gm(b, nb, ldk(a + a_split - 1)) + 1;
    wt.spawn(F_MERGE, Q_WAVE, kDummySlot, MergeArg{a, b, dst, a_split, b_split});
    wt.spawn(F_MERGE, Q_WAVE, kDummySlot, MergeArg{a + a_split, b + b_split, dst + a_split + b_split, na - a_split,
                                                    nb - b_split});
    if (wt.taskwait(1, Q_WAVE)) return;
    wt.finish();
}

// One chunk of a sort root's checksum: lane-strided terms, a wave sum, one atomic add.
template <class W>
__device__ void csum_task(W& wt) {
    const uint32_t* keys = wt.template arg<const uint32_t*>(0);
    const int64_t n = wt.template arg<int64_t>(8);
    const int64_t pos0 = wt.template arg<int64_t>(16);
    uint64_t* acc = wt.template arg<uint64_t*>(24);
    uint64_t part = 0;
    for (int64_t i = wt.lane; i < n; i += kWave) part += sort_mix((uint64_t)(pos0 + i), ldk(keys + i));
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) part += shfl(part, wt.lane ^ d);
    if (wt.lane == 0) atomic_add_noret(acc, part);
    wt.finish();
}

template <class W>
__device__ __forceinline__ void Program::run_wave(W& wt) {
    if (wt.func == F_MSORT) msort_task(wt);
    else if (wt.func == F_MERGE) merge_task(wt);
    else if (wt.func == F_CSUM) csum_task(wt);
    else {
        if (wt.lane == 0) GTAP_DEVICE_ERROR(*wt.dev, GTAP_ERROR_INVALID_FUNCTION, wt.tid, wt.queue, wt.func, F_CSUM);
        wt.finish();
    }
}

}  // namespace stream
}  // namespace gtap

using StreamProgram = gtap::stream::Program;
static_assert(StreamProgram::kFrameBytes + gtap::kHdrBytes <= gtap::host::kBuiltinStride, "frame exceeds stride");

// ======================================================================== host side

namespace {

using namespace gtap;

uint64_t splitmix64_next(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Task ids a job may hold at once, from its kind and size: at most the job's whole task
// count (a record is live from its spawn to its completion). fib(n) runs 2 F(n+1) - 1 tasks;
// n-queens(N) one per partial placement (counted here once per N); a sort job's msort tree
// (leaves of kWaveLeaf keys), its merge tasks (split down to merge_leaf keys, at most about four per
// merge_leaf outputs per level), its checksum tasks and its root.
uint64_t fib_tasks_host(int n) {
    uint64_t a = 0, b = 1;  // F(0), F(1)
    for (int i = 0; i < n + 1; ++i) {
        const uint64_t c = a + b;
        a = b;
        b = c;
    }
    return 2 * a - 1;  // a = F(n + 1)
}
uint64_t nq_nodes_rec(int n, int row, uint32_t cols, uint32_t diag, uint32_t anti) {
    if (row == n) return 1;
    uint64_t c = 1;
    uint32_t avail = ((1u << n) - 1u) & ~(cols | diag | anti);
    while (avail) {
        const uint32_t p = avail & (0u - avail);
        avail -= p;
        c += nq_nodes_rec(n, row + 1, cols | p, (diag | p) << 1, (anti | p) >> 1);
    }
    return c;
}
uint64_t nq_tasks_host(int n) {
    static uint64_t memo[17] = {0};
    if (n < 1 || n > 16) return 1;
    if (!memo[n]) memo[n] = nq_nodes_rec(n, 0, 0u, 0u, 0u);
    return memo[n];
}
uint64_t sort_tasks_host(int64_t n, int64_t merge_leaf) {
    const uint64_t leaves = (uint64_t)((n + stream::kWaveLeaf - 1) / stream::kWaveLeaf);
    uint64_t levels = 1;
    while ((1ull << levels) < leaves + 1) ++levels;
    const uint64_t merge_per_level = 4 * (uint64_t)((n + merge_leaf - 1) / merge_leaf);
    return 1 + 2 * leaves + levels * merge_per_level + (uint64_t)((n + stream::kCsumChunk - 1) / stream::kCsumChunk);
}

uint64_t host_now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

// Job i of the mixed stream from `seed` (40% fib n in [18,28], 30% n-queens N in [8,12],
// 30% mergesort of floor(2^(10+10u)) uint32 keys); kind 0 fib, 1 n-queens, 2 sort.
extern "C" int gtap_jobs_generate(uint64_t seed, int32_t njobs, int32_t* kind, int32_t* param, uint64_t* key_seed) {
    if (njobs < 0 || (njobs > 0 && (!kind || !param || !key_seed))) return host::invalid_argument("gtap_jobs_generate");
    for (int32_t i = 0; i < njobs; ++i) {
        uint64_t s = seed ^ ((uint64_t)(2 * (int64_t)i + 1) * 0xD1B54A32D192ED03ULL);
        const uint64_t r0 = splitmix64_next(s);
        const uint64_t r1 = splitmix64_next(s);
        const uint32_t k = (uint32_t)(r0 >> 32) % 10u;
        key_seed[i] = 0;
        if (k < 4) {
            kind[i] = 0;
            param[i] = 18 + (int32_t)((uint32_t)(r1 >> 32) % 11u);
        } else if (k < 7) {
            kind[i] = 1;
            param[i] = 8 + (int32_t)((uint32_t)(r1 >> 32) % 5u);
        } else {
            const double u = (double)(r1 >> 11) * (1.0 / 9007199254740992.0);
            double len = std::floor(std::exp2(10.0 + 10.0 * u));
            len = std::min(std::max(len, 1024.0), 1048576.0);
            kind[i] = 2;
            param[i] = (int32_t)len;
            key_seed[i] = r1 ^ 0x94D049BB133111EBULL;
        }
    }
    return 0;
}

extern "C" int gtap_gen_keys(uint64_t seed, uint64_t n, uint32_t* out);

struct gtap_stream {
    int32_t njobs = 0;
    std::vector<int32_t> kind, param;
    std::vector<uint64_t> kseed;
    std::vector<int64_t> key_off;   // offset of each sort job's keys (-1 otherwise)
    int64_t total_keys = 0;
    uint32_t* d_src = nullptr;      // pristine unsorted keys of every sort job
    uint32_t* d_work = nullptr;     // sorted in place
    uint32_t* d_tmp = nullptr;
    std::vector<gtap_job> table;    // job table template
    std::vector<uint64_t> tasks;    // task ids each job may hold at once (its reservation before capping)
    gtap_job* d_table = nullptr;    // the same, device-resident (source of the batched graph's copy)
    host::ThreadGraph graph;        // batched launch, captured on first use
    std::vector<double> exec_ms;    // last run: device time from admission to completion per job
    int32_t merge_leaf = stream::kWaveMergeLeaf;  // sort jobs: merges of at most this many keys on one wave
};

extern "C" int gtap_stream_destroy(gtap_stream* S);

extern "C" int gtap_stream_create(uint64_t seed, int32_t njobs, gtap_stream** out) {
    if (!out || njobs <= 0) return host::invalid_argument("gtap_stream_create");
    host::Runtime& R = host::rt();
    if (!R.initialized) {
        int st = gtap_initialize();
        if (st) return st;
    }
    if ((uint32_t)njobs > R.nanchors) return host::invalid_argument("gtap_stream_create: more jobs than max_jobs");
    gtap_stream* S = new gtap_stream();
    S->njobs = njobs;
    S->kind.resize(njobs);
    S->param.resize(njobs);
    S->kseed.resize(njobs);
    S->key_off.assign(njobs, -1);
    S->exec_ms.assign(njobs, 0.0);
    gtap_jobs_generate(seed, njobs, S->kind.data(), S->param.data(), S->kseed.data());
    for (int32_t i = 0; i < njobs; ++i) {
        if (S->kind[i] == 2) {
            S->key_off[i] = S->total_keys;
            S->total_keys += S->param[i];
        }
    }
    const size_t bytes = (size_t)std::max<int64_t>(S->total_keys, 1) * sizeof(uint32_t);
    hipError_t e = hipMalloc(&S->d_src, bytes);
    if (e == hipSuccess) e = hipMalloc(&S->d_work, bytes);
    if (e == hipSuccess) e = hipMalloc(&S->d_tmp, bytes);
    if (e != hipSuccess) {
        delete S;
        return host::hip_status(e, "gtap_stream_create alloc");
    }
    std::vector<uint32_t> keys((size_t)std::max<int64_t>(S->total_keys, 1));
    for (int32_t i = 0; i < njobs; ++i)
        if (S->kind[i] == 2) gtap_gen_keys(S->kseed[i], (uint64_t)S->param[i], keys.data() + S->key_off[i]);
    e = hipMemcpy(S->d_src, keys.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&S->d_table, (size_t)njobs * sizeof(gtap_job));
    if (e != hipSuccess) {
        gtap_stream_destroy(S);
        return host::hip_status(e, "gtap_stream_create upload");
    }
    S->table.resize(njobs);
    S->tasks.resize(njobs);
    for (int32_t i = 0; i < njobs; ++i) {
        S->tasks[i] = S->kind[i] == 0 ? fib_tasks_host(S->param[i])
                                       : (S->kind[i] == 1 ? nq_tasks_host(S->param[i])
                                                          : sort_tasks_host(S->param[i], S->merge_leaf));
        gtap_job j;
        memset(&j, 0, sizeof(j));
        j.status = GTAP_JOB_READY;
        j.kind = (uint32_t)S->kind[i];
        j.queue = stream::Q_LANE;
        if (S->kind[i] == 0) {
            j.func = stream::F_FIB;
            stream::FibArg a{S->param[i], 0};
            memcpy(j.args, &a, sizeof(a));
        } else if (S->kind[i] == 1) {
            j.func = stream::F_NQ;
            stream::NqArg a{0, 0u, 0u, 0u, S->param[i], 0};
            memcpy(j.args, &a, sizeof(a));
        } else {
            j.func = stream::F_SROOT;
            stream::SortRootArg a{S->d_work + S->key_off[i], S->d_tmp + S->key_off[i], (int64_t)S->param[i]};
            memcpy(j.args, &a, sizeof(a));
        }
        S->table[i] = j;
    }
    e = hipMemcpy(S->d_table, S->table.data(), (size_t)njobs * sizeof(gtap_job), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        gtap_stream_destroy(S);
        return host::hip_status(e, "gtap_stream_create upload");
    }
    *out = S;
    return 0;
}

extern "C" int gtap_stream_destroy(gtap_stream* S) {
    if (!S) return 0;
    host::destroy_graph(S->graph);
    (void)hipFree(S->d_table);
    (void)hipFree(S->d_src);
    (void)hipFree(S->d_work);
    (void)hipFree(S->d_tmp);
    delete S;
    return 0;
}

extern "C" int gtap_stream_info(gtap_stream* S, int32_t* kind, int32_t* param, uint64_t* key_seed, int64_t* total_keys) {
    if (!S) return host::invalid_argument("gtap_stream_info");
    if (kind) memcpy(kind, S->kind.data(), S->kind.size() * sizeof(int32_t));
    if (param) memcpy(param, S->param.data(), S->param.size() * sizeof(int32_t));
    if (key_seed) memcpy(key_seed, S->kseed.data(), S->kseed.size() * sizeof(uint64_t));
    if (total_keys) *total_keys = S->total_keys;
    return 0;
}

// Sort jobs merge runs of at most merge_leaf keys on one wave and split larger merges into
// merge tasks (the stream's counterpart of the reference's merge cutoff,
// evaluation/2-comparison/cilksort/gtap_cilksort.cu:12, 112). Changing it re-captures the batched
// graph and re-derives each sort job's reservation.
extern "C" int gtap_stream_set_merge_leaf(gtap_stream* S, int32_t merge_leaf) {
    if (!S || merge_leaf < stream::kWaveLeaf || merge_leaf > stream::kMaxWaveMergeLeaf)
        return host::invalid_argument("gtap_stream_set_merge_leaf: merge_leaf must be in [512, 16384]");
    if (merge_leaf == S->merge_leaf) return 0;
    host::destroy_graph(S->graph);
    S->merge_leaf = merge_leaf;
    for (int32_t i = 0; i < S->njobs; ++i)
        if (S->kind[i] == 2) S->tasks[i] = sort_tasks_host(S->param[i], merge_leaf);
    return 0;
}

// Copies the (sorted, after a run) keys of sort job `job` to out.
extern "C" int gtap_stream_sorted_keys(gtap_stream* S, int32_t job, uint32_t* out) {
    if (!S || job < 0 || job >= S->njobs || S->kind[job] != 2 || !out) return host::invalid_argument("gtap_stream_sorted_keys");
    GTAP_HIP_TRY(hipMemcpy(out, S->d_work + S->key_off[job], (size_t)S->param[job] * 4, hipMemcpyDeviceToHost));
    return 0;
}

static int stream_reset_keys(gtap_stream* S) {
    host::Runtime& R = host::rt();
    if (S->total_keys > 0)
        GTAP_HIP_TRY(hipMemcpyAsync(S->d_work, S->d_src, (size_t)S->total_keys * 4, hipMemcpyDeviceToDevice, R.stream));
    GTAP_HIP_TRY(hipStreamSynchronize(R.stream));
    return 0;
}

// All jobs in one persistent launch, replayed from a HIP graph captured on the first call
// (job-table copy, state reset, persistent kernel; re-captured when the runtime's buffers or
// configuration change). results[i] = root return value (sort jobs: the checksum of the sorted keys); latency_ms[i] =
// completion time of job i after the launch start (all jobs arrive at the start).
// stats->kernel_ms is the time of the whole graph.
extern "C" int gtap_stream_run_batched(gtap_stream* S, int64_t* results, double* latency_ms, gtap_run_stats* stats) {
    if (!S) return host::invalid_argument("gtap_stream_run_batched");
    host::Runtime& R = host::rt();
    int st = stream_reset_keys(S);
    if (st) return st;
    st = host::capture_thread<StreamProgram>(S->graph, StreamProgram::Args{S->merge_leaf}, GTAP_TERM_JOBS, (uint32_t)S->njobs,
                                             S->d_table, (uint32_t)S->njobs);
    if (st) return st;
    st = host::replay_thread(S->graph, stats);
    if (st) return st;
    std::vector<gtap_job> out(S->njobs);
    GTAP_HIP_TRY(hipMemcpy(out.data(), R.jobs, out.size() * sizeof(gtap_job), hipMemcpyDeviceToHost));
    uint64_t t0 = UINT64_MAX;
    for (auto& j : out) t0 = std::min(t0, j.t_start);
    for (int32_t i = 0; i < S->njobs; ++i) {
        if (out[i].status != GTAP_JOB_DONE) return host::fail(GTAP_ERROR_INVALID_TASKWAIT, "stream job did not complete");
        if (results) results[i] = out[i].result;
        if (latency_ms) latency_ms[i] = (double)(out[i].t_end - t0) * 1e-5;  // 100 MHz ticks
        S->exec_ms[i] = (double)(out[i].t_end - out[i].t_start) * 1e-5;
    }
    return 0;
}

// The plain form: one persistent launch per job, in job order. latency_ms[i] = host time
// from the start of the sequence to the end of job i's launch; *total_ms = whole sequence.
extern "C" int gtap_stream_run_per_job(gtap_stream* S, int64_t* results, double* latency_ms, double* total_ms) {
    if (!S) return host::invalid_argument("gtap_stream_run_per_job");
    host::Runtime& R = host::rt();
    int st = stream_reset_keys(S);
    if (st) return st;
    const uint64_t t0 = host_now_ns();
    for (int32_t i = 0; i < S->njobs; ++i) {
        int64_t r = 0;
        gtap_job out;
        st = host::run_single_job<StreamProgram>(StreamProgram::Args{S->merge_leaf}, S->table[i], &r, nullptr, GTAP_TERM_JOBS, &out);
        if (st) return st;
        S->exec_ms[i] = (double)(out.t_end - out.t_start) * 1e-5;
        if (results) results[i] = r;
        if (latency_ms) latency_ms[i] = (double)(host_now_ns() - t0) * 1e-6;
    }
    if (total_ms) *total_ms = (double)(host_now_ns() - t0) * 1e-6;
    (void)R;
    return 0;
}

// Device time from admission to completion of each job in the last run (batched, per job or
// served), in ms.
extern "C" int gtap_stream_exec_ms(gtap_stream* S, double* exec_ms) {
    if (!S || !exec_ms) return host::invalid_argument("gtap_stream_exec_ms");
    memcpy(exec_ms, S->exec_ms.data(), S->exec_ms.size() * sizeof(double));
    return 0;
}

// Poisson arrivals at `rate` jobs/s (exponential gaps from arrival_seed). Resident form
// (per_job_launch = 0): one served launch; the host loop writes each job into the pinned
// job table at its arrival time and publishes the count, the kernel admits jobs as they
// appear (up to max_live concurrently running), the host polls completions. Plain form
// (per_job_launch = 1): each job is launched on its own when it has arrived and the
// previous launch has finished. latency_ms[i] = completion seen by the host - arrival.
extern "C" int gtap_stream_serve(gtap_stream* S, double rate, uint64_t arrival_seed, int32_t per_job_launch,
                                 int32_t max_live, int64_t* results, double* latency_ms, double* total_ms,
                                 gtap_run_stats* stats) {
    if (!S || !(rate > 0)) return host::invalid_argument("gtap_stream_serve");
    host::Runtime& R = host::rt();
    int st = stream_reset_keys(S);
    if (st) return st;
    const int32_t n = S->njobs;
    std::vector<uint64_t> arrival(n);
    {
        uint64_t s = arrival_seed ^ 0xA0761D6478BD642FULL;
        double t = 0;
        for (int32_t i = 0; i < n; ++i) {
            const double u = ((double)(splitmix64_next(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            t += -std::log(u) / rate;
            arrival[i] = (uint64_t)(t * 1e9);
        }
    }
    std::vector<double> lat(n, 0.0);
    std::vector<int64_t> res(n, 0);
    if (per_job_launch) {
        const uint64_t t0 = host_now_ns();
        for (int32_t i = 0; i < n; ++i) {
            while (host_now_ns() - t0 < arrival[i]) std::this_thread::yield();
            int64_t r = 0;
            gtap_job out;
            st = host::run_single_job<StreamProgram>(StreamProgram::Args{S->merge_leaf}, S->table[i], &r, nullptr, GTAP_TERM_JOBS, &out);
            if (st) return st;
            S->exec_ms[i] = (double)(out.t_end - out.t_start) * 1e-5;
            res[i] = r;
            lat[i] = (double)(host_now_ns() - t0 - arrival[i]) * 1e-6;
        }
        if (total_ms) *total_ms = (double)(host_now_ns() - t0) * 1e-6;
    } else {
        if (!R.jobs_host) {
            GTAP_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&R.jobs_host), (size_t)R.nanchors * sizeof(gtap_job),
                                       hipHostMallocMapped | hipHostMallocCoherent));
            GTAP_HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&R.jobs_host_dev), R.jobs_host, 0));
        }
        // The job table is a ring of `ring` slots: job i uses slot i % ring, and the host writes
        // job i only after job i - ring has completed and been collected, so a served stream
        // is not bounded by the table.
        const uint32_t live_cap = max_live > 0 ? (uint32_t)max_live : 1024u;
        uint32_t ring = std::max<uint32_t>(2 * live_cap, 256u);
        ring = std::min<uint32_t>(ring, std::min<uint32_t>((uint32_t)n, R.nanchors));
        volatile gtap_job* tab = R.jobs_host;
        volatile int32_t* hctl = R.host_ctl;
        hctl[HCTL_JOB_AVAIL * 32] = 0;
        hctl[HCTL_STOP * 32] = 0;
        __atomic_thread_fence(__ATOMIC_SEQ_CST);
        // the launch itself (asynchronous): served mode, admission by live jobs and free ids
        // the timeline-profiling build of the kernel when profiling is on
        const bool prof = R.prof_max != 0;
        static uint32_t max_blocks[2] = {0, 0};
        const void* kfn = prof ? reinterpret_cast<const void*>(&thread_kernel<StreamProgram, kSchedWS, true>)
                               : reinterpret_cast<const void*>(&thread_kernel<StreamProgram, kSchedWS, false>);
        if (max_blocks[prof] == 0) {
            st = host::occupancy_blocks(kfn, 0, &max_blocks[prof]);
            if (st) return st;
        }
        uint32_t blocks = max_blocks[prof];
        if (R.cfg.blocks && R.cfg.blocks < blocks) blocks = R.cfg.blocks;
        if (blocks * 4 > R.max_workers) blocks = R.max_workers / 4;
        GtapDev dev = host::make_dev(blocks * 4, GTAP_TERM_SERVE, (uint32_t)n, R.jobs_host_dev, ring);
        dev.max_live_jobs = live_cap;
        // Admission by free task ids. A wave may park ids it freed in its LDS stack (fewer than
        // kFreeCap) and in its private HBM stack (priv_cap, a quarter of the pool spread over the
        // waves), where other waves cannot take them; the rest of the pool is the budget that
        // the reservations of the running jobs share. A job reserves the task ids it may hold
        // at once (its task count, at most half the budget) and releases them when it
        // completes; a job whose reservation does not fit waits in the job table. With the
        // default pool every job's task count fits under the cap, so the pool cannot run dry.
        {
            uint64_t pc = (uint64_t)R.nrecs / (4ull * dev.nworkers) / kChunk * kChunk;
            dev.priv_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(pc, kChunk), kPrivCap);
            const int64_t parked = (int64_t)dev.nworkers * (int64_t)(kFreeCap + dev.priv_cap);
            const int64_t budget = (int64_t)R.nrecs - parked;
            if (budget < (int64_t)(64 * kChunk))
                return host::invalid_argument("gtap_stream_serve: the task-id pool is too small for the served workers");
            dev.admit_budget = (uint32_t)std::min<int64_t>(budget / kChunk, 0x7FFFFFFF);
            dev.job_units = R.job_units;
        }
        const uint64_t cap_ids = (uint64_t)dev.admit_budget * kChunk / 2;
        auto reserve_of = [&](int32_t i) { return std::min<uint64_t>(S->tasks[i], cap_ids); };
        const double expected_s = arrival.back() * 1e-9;
        const uint64_t need = (uint64_t)((expected_s * 3.0 + 30.0) * 1e8);
        if (dev.timeout_ticks < need) dev.timeout_ticks = need;
        st = host::reset_launch_state(dev, R.stream);
        if (st) return st;
        GTAP_HIP_TRY(hipEventRecord(R.ev0, R.stream));
        if (prof)
            hipLaunchKernelGGL((thread_kernel<StreamProgram, kSchedWS, true>), dim3(blocks), dim3(256), 0, R.stream, dev,
                               StreamProgram::Args{S->merge_leaf});
        else
            hipLaunchKernelGGL((thread_kernel<StreamProgram, kSchedWS, false>), dim3(blocks), dim3(256), 0, R.stream, dev,
                               StreamProgram::Args{S->merge_leaf});
        GTAP_HIP_TRY(hipGetLastError());
        GTAP_HIP_TRY(hipEventRecord(R.ev1, R.stream));
        st = host::enqueue_summary(dev, R.stream, false);
        if (st) return st;
        const uint64_t t0 = host_now_ns();
        int32_t submitted = 0, done = 0, first_pending = 0;
        std::vector<char> seen(n, 0);
        const uint64_t deadline = t0 + (uint64_t)((expected_s * 3.0 + 60.0) * 1e9);
        while (done < n) {
            const uint64_t now = host_now_ns();
            // submit the arrived jobs whose ring slot is free (its previous job collected)
            while (submitted < n && now - t0 >= arrival[submitted] &&
                   (submitted < (int32_t)ring || seen[submitted - (int32_t)ring])) {
                gtap_job j = S->table[submitted];
                j.t_submit = now;
                j.reserve_ids = reserve_of(submitted);
                j.status = GTAP_JOB_EMPTY;
                volatile gtap_job* slot = &tab[submitted % (int32_t)ring];
                memcpy((void*)slot, &j, sizeof(j));
                __atomic_thread_fence(__ATOMIC_SEQ_CST);
                slot->status = GTAP_JOB_READY;
                ++submitted;
                __atomic_thread_fence(__ATOMIC_SEQ_CST);
                hctl[HCTL_JOB_AVAIL * 32] = submitted;
            }
            // completions: only the jobs in flight, oldest first
            while (first_pending < submitted && seen[first_pending]) ++first_pending;
            for (int32_t i = first_pending; i < submitted; ++i) {
                volatile gtap_job* slot = &tab[i % (int32_t)ring];
                if (!seen[i] && slot->status == GTAP_JOB_DONE) {
                    seen[i] = 1;
                    ++done;
                    const uint64_t tn = host_now_ns();
                    lat[i] = (double)(tn - t0 - arrival[i]) * 1e-6;
                    res[i] = slot->result;
                    S->exec_ms[i] = (double)(slot->t_end - slot->t_start) * 1e-5;
                }
            }
            if (hipStreamQuery(R.stream) == hipSuccess) break;  // the kernel ended (error or done)
            if (host_now_ns() > deadline) break;
        }
        hctl[HCTL_STOP * 32] = 1;
        __atomic_thread_fence(__ATOMIC_SEQ_CST);
        GTAP_HIP_TRY(hipStreamSynchronize(R.stream));
        if (total_ms) *total_ms = (double)(host_now_ns() - t0) * 1e-6;
        st = host::finish_launch(dev, blocks, stats);
        if (st) return st;
        for (int32_t i = 0; i < n; ++i) {
            if (seen[i]) continue;
            // not collected in the loop: still in its slot (job i + ring was never written)
            if (i >= submitted) return host::fail(GTAP_ERROR_INVALID_TASKWAIT, "served job was never submitted");
            volatile gtap_job* slot = &tab[i % (int32_t)ring];
            if (slot->status != GTAP_JOB_DONE) return host::fail(GTAP_ERROR_INVALID_TASKWAIT, "served job did not complete");
            seen[i] = 1;
            lat[i] = (double)(host_now_ns() - t0 - arrival[i]) * 1e-6;
            res[i] = slot->result;
            S->exec_ms[i] = (double)(slot->t_end - slot->t_start) * 1e-5;
        }
    }
    if (results) memcpy(results, res.data(), n * sizeof(int64_t));
    if (latency_ms) memcpy(latency_ms, lat.data(), n * sizeof(double));
    return 0;
}
