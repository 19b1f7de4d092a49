// prog_cilksort.hip -- cilksort of uint32 keys as a block-cooperative task program.
//
// Task structure of the reference's evaluation/2-comparison/cilksort/gtap_cilksort.cu:89-167:
//
//   cilksort(arr, n, tmp):  n <= LEAF        -> the workgroup sorts arr[0..n) in LDS
//                           else four-way split len12 = n/2, len1 = len12/2, len2 = len12-len1,
//                           len34 = n-len12, len3 = len34/2, len4 = len34-len3;
//                           4 x task cilksort; taskwait;
//                           task cilkmerge(arr,len1, arr+len1,len2, tmp);
//                           task cilkmerge(arr+len12,len3, arr+len12+len3,len4, tmp+len12); taskwait;
//                           task cilkmerge(tmp,len12, tmp+len12,len34, arr); taskwait
//   cilkmerge(a,na,b,nb,dst): make na >= nb; nb == 0 -> copy;
//                           na + nb <= MERGE_LEAF -> the workgroup merges in LDS tiles;
//                           reference: a_split = (na+1)/2,
//                                b_split = binary_search(b, nb, a[a_split-1]) + 1   (low = -1, :40-50)
//                           task cilkmerge(a,a_split, b,b_split, dst);
//                           task cilkmerge(a+a_split,na-a_split, b+b_split,nb-b_split,
//                                          dst+a_split+b_split); taskwait
//                           here: up to 64 chunks of equal output length, their boundaries found by
//                           merge-path searches run by the workgroup's threads in parallel, each
//                           chunk a cilkmerge task; taskwait (the output is the same sorted run)
//
// Differences from the reference's thread-mode program: uint32 keys; the leaf is `n <= 4096`
// keys sorted by a 256-thread workgroup (bitonic network in LDS) instead of `n < 64` keys by
// one thread; merges below MERGE_LEAF keys are done by a workgroup with merge-path tiles
// (4096 outputs per tile, staged through LDS) instead of one thread. The output is the
// unique sorted permutation of the input whatever the schedule.
//
// All key traffic uses sc1 (agent-coherent) accesses: a merge may run on another XCD than
// the tasks that wrote its inputs.
#include "../device/gtap_block.hpp"
#include "../host/gtap_launch.hpp"

namespace gtap {
namespace cilk {

enum : uint32_t { F_SORT = 1, F_MERGE = 2 };
constexpr int kTile = 4096;
constexpr int kMaxFan = 64;  // merge chunks spawned by one large merge (<= kBlockMaxSpawn)
constexpr uint32_t kT = kBlockThreads;

struct SortArg {
    uint32_t* arr;
    uint32_t* tmp;
    int64_t n;
};
struct MergeArg {
    const uint32_t* a;
    const uint32_t* b;
    uint32_t* dst;
    int64_t na;
    int64_t nb;
};
struct CilkArgs {
    int64_t leaf;        // <= 4096
    int64_t merge_leaf;  // merges of at most this many keys are done by one workgroup
};

struct Program {
    static constexpr int kWavesPerSimd = 4;  // four 256-thread workgroups per CU
    static constexpr int kFrameWords = 10;
    static constexpr uint32_t kMaxFunc = F_MERGE;
    static constexpr uint32_t kFrameBytes = 40;
    using Args = CilkArgs;
    struct Shared {
        uint32_t keys[kTile];
        uint32_t out[kTile];
        int64_t split[kMaxFan + 2];
    };
    template <class B>
    __device__ __forceinline__ static void run(B& blk);
};

// keys live in device memory (the caller's arrays and the scratch buffer)
__device__ __forceinline__ uint32_t ldk(const uint32_t* p) { return gld32(p); }
__device__ __forceinline__ void stk(uint32_t* p, uint32_t v) { gst32(p, v); }

// gtap_cilksort.cu:40-50: largest index with x[index] <= val, or -1.
__device__ __forceinline__ int64_t binary_search_gm(const uint32_t* x, int64_t len, uint32_t val) {
    int64_t low = -1, high = len;
    while (low + 1 < high) {
        const int64_t mid = (low + high) / 2;
        if (ldk(x + mid) <= val) low = mid;
        else high = mid;
    }
    return low;
}

// Per-thread batch of kPer keys of a kTile-key tile: element e = k * kT + t (k < kPer), so the
// loads of a warp are contiguous, and all kPer loads are issued before any is used.
constexpr int kPer = kTile / (int)kT;  // 16

// Compare-exchange of a pair held by one thread: afterwards x <= y when `up`, else x >= y.
__device__ __forceinline__ void cmpx(uint32_t& x, uint32_t& y, bool up) {
    const uint32_t lo = x < y ? x : y, hi = x < y ? y : x;
    x = up ? lo : hi;
    y = up ? hi : lo;
}

// Partner exchange of a bitonic step whose partner slot is in lane (lane ^ m) of the same
// wave: every register of the thread meets the same register of that lane.
__device__ __forceinline__ uint32_t xlane(uint32_t v, int m) { return (uint32_t)__shfl_xor((int)v, m, kWave); }

// Workgroup sort of arr[0..n), n <= 4096: bitonic sort of 4096 slots (padded with 0xFFFFFFFF,
// which sorts to the end and is never written back). Thread t owns slots 16t .. 16t+15 in
// registers. A step with partner distance j pairs slot s with s ^ j:
//   j < 16          the partner is in the same thread: registers;
//   16 <= j < 1024  the partner is register (s & 15) of lane (lane ^ j/16) of the same wave:
//                   a cross-lane exchange, no barrier;
//   j >= 1024       the partner is in another wave: through LDS with barriers (3 steps).
// Barriers here order LDS only (lds_barrier): the global loads and stores of a task need no
// ordering among its threads, and the scheduler's barrier after the task waits for them.
template <class B>
__device__ void leaf_sort(B& blk, uint32_t* arr, int64_t n) {
    uint32_t* k = blk.sh->keys;
    const uint32_t t = threadIdx.x;
    {
        uint32_t r[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const uint32_t e = (uint32_t)q * kT + t;
            r[q] = (int64_t)e < n ? ldk(arr + e) : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q) k[(uint32_t)q * kT + t] = r[q];
    }
    lds_barrier();
    uint32_t x[16];
    const uint32_t g0 = 16u * t;  // global slot of x[0]
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = k[g0 + (uint32_t)i];
    // sizes 2..16: entirely inside each thread's 16 slots
#pragma unroll
    for (uint32_t size = 2; size <= 16; size <<= 1) {
#pragma unroll
        for (uint32_t j = size >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (uint32_t i = 0; i < 16; ++i) {
                if ((i & j) == 0) cmpx(x[i], x[i + j], ((g0 + i) & size) == 0);
            }
        }
    }
    for (uint32_t size = 32; size <= (uint32_t)kTile; size <<= 1) {
        const bool up = (g0 & size) == 0;  // direction of this thread's slots at this size
        for (uint32_t j = size >> 1; j >= 16; j >>= 1) {
            const uint32_t m = j / 16;  // partner thread = t ^ m
            const bool low = (t & m) == 0;
            const bool keep_min = (low == up);
            if (j >= 1024) {
                lds_barrier();  // every thread has read the previous contents
#pragma unroll
                for (int i = 0; i < 16; ++i) k[g0 + (uint32_t)i] = x[i];
                lds_barrier();
                const uint32_t p0 = 16u * (t ^ m);
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const uint32_t y = k[p0 + (uint32_t)i];
                    x[i] = keep_min ? (x[i] < y ? x[i] : y) : (x[i] < y ? y : x[i]);
                }
            } else {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const uint32_t y = xlane(x[i], (int)m);
                    x[i] = keep_min ? (x[i] < y ? x[i] : y) : (x[i] < y ? y : x[i]);
                }
            }
        }
#pragma unroll
        for (uint32_t j = 8; j > 0; j >>= 1) {
#pragma unroll
            for (uint32_t i = 0; i < 16; ++i) {
                if ((i & j) == 0) cmpx(x[i], x[i + j], up);
            }
        }
    }
    lds_barrier();
#pragma unroll
    for (int i = 0; i < 16; ++i) k[g0 + (uint32_t)i] = x[i];
    lds_barrier();
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const uint32_t e = (uint32_t)q * kT + t;
        if ((int64_t)e < n) stk(arr + e, k[e]);
    }
}

// Workgroup copy of n keys, kTile keys per step with all loads of a step in flight.
template <class B>
__device__ void block_copy(B& blk, const uint32_t* src, uint32_t* dst, int64_t n) {
    (void)blk;
    const uint32_t t = threadIdx.x;
    for (int64_t base = 0; base < n; base += kTile) {
        uint32_t r[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int64_t e = base + (int64_t)q * kT + t;
            r[q] = e < n ? ldk(src + e) : 0u;
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int64_t e = base + (int64_t)q * kT + t;
            if (e < n) stk(dst + e, r[q]);
        }
    }
}

// Merge-path splits of nd diagonals at once (nd <= kT), d_k = min(d0 + k * step, na + nb):
// split[k] = number of a elements among the first d_k outputs of merge(a, b), a first on
// ties. The result is that of merge_path; the search is k-ary instead of binary. The
// workgroup's threads form groups of G lanes (G a power of two, G * nd <= kT, G <= 64, so a
// group never straddles a wave); each round a group evaluates the predicate
// a[i] <= b[d - 1 - i] at G points of its range together (one load pair per lane) and keeps
// the sub-range between the last true and the first false probe, so a range of R candidates
// takes about log_(G+1) R dependent global loads instead of log_2 R.
template <class LDA, class LDB>
__device__ void block_split_search(LDA lda, int64_t na, LDB ldb, int64_t nb, int64_t d0, int64_t step, int nd,
                                   int64_t* split) {
    const uint32_t t = threadIdx.x;
    int G = kWave;
    while (G > 1 && G * nd > (int)kT) G >>= 1;
    const int grp = (int)t / G, k = (int)t % G;
    const int64_t total = na + nb;
    const bool mine = grp < nd;
    int64_t d = 0, lo = 0, hi = 0;
    if (mine) {
        d = d0 + (int64_t)grp * step;
        if (d > total) d = total;
        lo = d > nb ? d - nb : 0;
        hi = d < na ? d : na;
    }
    const uint32_t shift = (uint32_t)(lane_id() / G) * (uint32_t)G;  // this group's bits in a ballot
    const uint64_t gmask = G == kWave ? ~0ull : ((1ull << G) - 1ull);
    while (true) {
        const bool active = mine && lo < hi;
        if (ballot(active) == 0) break;
        const int64_t span = hi - lo;
        const bool fine = span <= G;  // every candidate gets a probe
        const int64_t ik = fine ? lo + k : lo + ((int64_t)(k + 1) * span) / (G + 1);
        bool p = false;
        if (active && (!fine || k < span)) p = lda(ik) <= ldb(d - 1 - ik);
        const int c = popc((ballot(p) >> shift) & gmask);  // the predicate holds on the first c probes
        if (active) {
            if (fine) {
                lo += c;
                hi = lo;
            } else {
                const int64_t nlo = c > 0 ? lo + ((int64_t)c * span) / (G + 1) + 1 : lo;
                const int64_t nhi = c < G ? lo + ((int64_t)(c + 1) * span) / (G + 1) : hi;
                lo = nlo;
                hi = nhi;
            }
        }
    }
    if (mine && k == 0) split[grp] = lo;
    lds_barrier();
}

// The inputs of one merge tile: a[i0 .. i0+la) then b[j0 .. j0+lb), loaded kPer per thread.
__device__ __forceinline__ void load_tile(const uint32_t* a, int64_t i0, int64_t la, const uint32_t* b, int64_t j0,
                                          int64_t lb, uint32_t (&r)[kPer]) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int64_t e = (int64_t)q * kT + t;
        r[q] = e < la ? ldk(a + i0 + e) : (e < la + lb ? ldk(b + j0 + (e - la)) : 0u);
    }
}

// Number of A elements among the first d outputs of merge(A, B) over LDS, A first on ties.
__device__ __forceinline__ int32_t merge_path_lds(const uint32_t* A, int32_t la, const uint32_t* Bv, int32_t lb,
                                                  int32_t d) {
    int32_t lo = d > lb ? d - lb : 0;
    int32_t hi = d < la ? d : la;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (A[mid] <= Bv[d - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Workgroup merge of a[0..na) and b[0..nb) into dst, tile by tile (merge path). The inputs of
// the next tile are loaded into registers while the current tile is merged in LDS. Each
// thread merges 16 consecutive outputs of a tile, keeping the current head of each input in
// a register, so each output costs one LDS load.
template <class B>
__device__ void leaf_merge(B& blk, const uint32_t* a, int64_t na, const uint32_t* b, int64_t nb, uint32_t* dst) {
    typename Program::Shared* sh = blk.sh;
    const uint32_t t = threadIdx.x;
    const int64_t total = na + nb;
    auto ga = [a](int64_t i) { return ldk(a + i); };
    auto gb = [b](int64_t i) { return ldk(b + i); };
    int64_t carry = 0;  // split at the first diagonal of the group
    for (int64_t base = 0; base < total; base += (int64_t)kTile * 16) {
        // splits at the ends of up to 16 tiles
        const int64_t ntiles_here = ((total - base) + kTile - 1) / kTile;
        const int ntl = ntiles_here < 16 ? (int)ntiles_here : 16;
        if (t == 0) sh->split[0] = carry;
        block_split_search(ga, na, gb, nb, base + kTile, kTile, ntl, sh->split + 1);
        carry = sh->split[ntl];
        // The splits of sorted inputs are monotone along both inputs; anything else (unsorted
        // runs) would index outside the tile, so it is reported and the merge abandoned.
        bool ok = true;
        for (int ti = 0; ti < ntl; ++ti) {
            const int64_t d0 = base + (int64_t)ti * kTile, d1 = (d0 + kTile < total) ? d0 + kTile : total;
            const int64_t i0 = sh->split[ti], i1 = sh->split[ti + 1];
            const int64_t la = i1 - i0, lb = (d1 - i1) - (d0 - i0);
            ok = ok && la >= 0 && lb >= 0 && i1 <= na && (d1 - i1) <= nb;
        }
        if (!ok) {
            if (t == 0) GTAP_DEVICE_ERROR(*blk.dev, GTAP_ERROR_INVALID_ARGUMENT, blk.tid, -1, (int)na, (int)nb);
            lds_barrier();
            return;
        }
        uint32_t r[kPer];
        {
            const int64_t i0 = sh->split[0], i1 = sh->split[1];
            const int64_t d0 = base, d1 = (d0 + kTile < total) ? d0 + kTile : total;
            load_tile(a, i0, i1 - i0, b, d0 - i0, (d1 - i1) - (d0 - i0), r);
        }
        for (int ti = 0; ti < ntl; ++ti) {
            const int64_t d0 = base + (int64_t)ti * kTile;
            const int64_t d1 = (d0 + kTile < total) ? d0 + kTile : total;
            const int64_t i0 = sh->split[ti], i1 = sh->split[ti + 1];
            const int32_t la = (int32_t)(i1 - i0), lb = (int32_t)((d1 - i1) - (d0 - i0));
#pragma unroll
            for (int q = 0; q < kPer; ++q) sh->keys[(uint32_t)q * kT + t] = r[q];
            lds_barrier();
            if (ti + 1 < ntl) {  // next tile's inputs in flight during this merge
                const int64_t e0 = d1, e1 = (e0 + kTile < total) ? e0 + kTile : total;
                const int64_t n0 = sh->split[ti + 1], n1 = sh->split[ti + 2];
                load_tile(a, n0, n1 - n0, b, e0 - n0, (e1 - n1) - (e0 - n0), r);
            }
            const uint32_t* K = sh->keys;  // A = K[0..la), B = K[la..la+lb)
            const int32_t len = (int32_t)(d1 - d0);
            const int32_t od = (int32_t)t * kPer;
            if (od < len) {
                int32_t i = merge_path_lds(K, la, K + la, lb, od);
                int32_t j = od - i;
                uint32_t av = i < la ? K[i] : 0u;
                uint32_t bv = j < lb ? K[la + j] : 0u;
                const int32_t cnt = (od + kPer < len) ? kPer : len - od;
                uint32_t outv[kPer];
#pragma unroll
                for (int o = 0; o < kPer; ++o) {
                    const bool takeA = i < la && (j >= lb || av <= bv);
                    outv[o] = takeA ? av : bv;
                    i += takeA ? 1 : 0;
                    j += takeA ? 0 : 1;
                    // the next head of the input just consumed
                    const int32_t at = takeA ? i : la + j;
                    const bool in = takeA ? i < la : j < lb;
                    const uint32_t v = (in && o + 1 < cnt) ? K[at] : 0u;
                    av = takeA ? v : av;
                    bv = takeA ? bv : v;
                }
                if (cnt == kPer) {
#pragma unroll
                    for (int o = 0; o < kPer; o += 4)
                        *reinterpret_cast<uint4*>(&sh->out[od + o]) = make_uint4(outv[o], outv[o + 1], outv[o + 2], outv[o + 3]);
                } else {
#pragma unroll
                    for (int o = 0; o < kPer; ++o)
                        if (o < cnt) sh->out[od + o] = outv[o];
                }
            }
            lds_barrier();
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const int32_t o = q * (int32_t)kT + (int32_t)t;
                if (o < len) stk(dst + d0 + o, sh->out[o]);
            }
        }
        lds_barrier();
    }
}

template <class B>
__device__ void sort_task(B& blk) {
    const CilkArgs& g = *blk.args;
    uint32_t* arr = blk.template arg<uint32_t*>(0);
    uint32_t* tmp = blk.template arg<uint32_t*>(8);
    const int64_t n = blk.template arg<int64_t>(16);
    const int64_t len12 = n / 2, len1 = len12 / 2, len2 = len12 - len1;
    const int64_t len34 = n - len12, len3 = len34 / 2, len4 = len34 - len3;
    switch (blk.get_state()) {
        case 0:
            if (n <= g.leaf) {
                leaf_sort(blk, arr, n);
                blk.finish();
                return;
            }
            if (threadIdx.x == 0) {
                blk.spawn(F_SORT, 0, SortArg{arr, tmp, len1});
                blk.spawn(F_SORT, 0, SortArg{arr + len1, tmp + len1, len2});
                blk.spawn(F_SORT, 0, SortArg{arr + len12, tmp + len12, len3});
                blk.spawn(F_SORT, 0, SortArg{arr + len12 + len3, tmp + len12 + len3, len4});
            }
            if (blk.taskwait(1)) return;
            [[fallthrough]];
        case 1:
            if (threadIdx.x == 0) {
                blk.spawn(F_MERGE, 0, MergeArg{arr, arr + len1, tmp, len1, len2});
                blk.spawn(F_MERGE, 0, MergeArg{arr + len12, arr + len12 + len3, tmp + len12, len3, len4});
            }
            if (blk.taskwait(2)) return;
            [[fallthrough]];
        case 2:
            if (threadIdx.x == 0) blk.spawn(F_MERGE, 0, MergeArg{tmp, tmp + len12, arr, len12, len34});
            if (blk.taskwait(3)) return;
            [[fallthrough]];
        case 3:
            blk.finish();
            return;
        default:
            if (threadIdx.x == 0) GTAP_DEVICE_ERROR(*blk.dev, GTAP_ERROR_INVALID_TASKWAIT, blk.tid, 0, blk.state, 3);
            blk.finish();
            return;
    }
}

template <class B>
__device__ void merge_task(B& blk) {
    const CilkArgs& g = *blk.args;
    const uint32_t* a = blk.template arg<const uint32_t*>(0);
    const uint32_t* b = blk.template arg<const uint32_t*>(8);
    uint32_t* dst = blk.template arg<uint32_t*>(16);
    int64_t na = blk.template arg<int64_t>(24);
    int64_t nb = blk.template arg<int64_t>(32);
    if (blk.get_state() != 0) {
        blk.finish();
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
        block_copy(blk, a, dst, na);
        blk.finish();
        return;
    }
    if (na + nb <= g.merge_leaf) {
        leaf_merge(blk, a, na, b, nb, dst);
        blk.finish();
        return;
    }
    // A large merge is cut into up to kMaxFan chunks of equal output length by merge-path
    // searches run in parallel (one per chunk boundary), and the chunks are spawned as merge
    // tasks: one search latency per level instead of one per halving.
    {
        typename Program::Shared* sh = blk.sh;
        const int64_t total = na + nb;
        int64_t fan = (total + g.merge_leaf - 1) / g.merge_leaf;
        if (fan > kMaxFan) fan = kMaxFan;
        if (fan < 2) fan = 2;
        const int64_t chunk = (total + fan - 1) / fan;
        auto ga = [a](int64_t i) { return ldk(a + i); };
        auto gb = [b](int64_t i) { return ldk(b + i); };
        const int64_t t = threadIdx.x;
        if (t == 0) {
            sh->split[0] = 0;
            sh->split[fan] = na;
        }
        block_split_search(ga, na, gb, nb, chunk, chunk, (int)fan - 1, sh->split + 1);
        if (t < fan) {
            int64_t d0 = t * chunk, d1 = d0 + chunk;
            if (d0 > total) d0 = total;
            if (d1 > total) d1 = total;
            const int64_t i0 = sh->split[t], i1 = sh->split[t + 1];
            const int64_t ca = i1 - i0, cb = (d1 - i1) - (d0 - i0);
            if (ca < 0 || cb < 0 || i1 > na || d1 - i1 > nb) {
                GTAP_DEVICE_ERROR(*blk.dev, GTAP_ERROR_INVALID_ARGUMENT, blk.tid, -1, (int)na, (int)nb);
            } else if (ca + cb > 0) {
                blk.spawn(F_MERGE, 0, MergeArg{a + i0, b + (d0 - i0), dst + d0, ca, cb});
            }
        }
    }
    if (blk.taskwait(1)) return;
    blk.finish();
}

template <class B>
__device__ __forceinline__ void Program::run(B& blk) {
    if (blk.func == F_SORT) sort_task(blk);
    else merge_task(blk);
}

}  // namespace cilk
}  // namespace gtap

using CilkProgram = gtap::cilk::Program;
static_assert(CilkProgram::kFrameBytes + gtap::kHdrBytes <= gtap::host::kBuiltinStride, "frame exceeds stride");
static_assert(CilkProgram::kFrameWords <= 16, "block frame cache holds 16 words");

// Sorts n uint32 keys in place (device buffers keys and tmp, n elements each).
// leaf <= 4096 keys are sorted by one workgroup in LDS (0 -> 4096); merges of at most
// merge_leaf keys are done by one workgroup (0 -> 32768).
// Replaces the program of evaluation/2-comparison/cilksort/gtap_cilksort.cu:169-281.
extern "C" int gtap_cilksort(uint32_t* keys, uint32_t* tmp, int64_t n, int64_t leaf, int64_t merge_leaf,
                             gtap_run_stats* stats) {
    using namespace gtap;
    if (!keys || !tmp || n < 0) return host::invalid_argument("gtap_cilksort: bad arguments");
    if (leaf <= 0) leaf = 4096;
    if (merge_leaf <= 0) merge_leaf = 32768;
    if (leaf > cilk::kTile) return host::invalid_argument("gtap_cilksort: leaf must be <= 4096");
    if (merge_leaf < 3) merge_leaf = 3;
    if (n == 0) {
        if (stats) memset(stats, 0, sizeof(*stats));
        return 0;
    }
    cilk::SortArg root{keys, tmp, n};
    gtap_job job = host::make_job(cilk::F_SORT, 0, &root, sizeof(root));
    return host::run_single_block_job<CilkProgram>(cilk::CilkArgs{leaf, merge_leaf}, job, nullptr, stats);
}

namespace gtap {
namespace cilk {

// Workspace of gtap_cilksort_noalloc: the merge buffer (`tmp` of gtap_cilksort) as a
// file-scope device array, so the call allocates nothing. Every GPU the library runs on holds
// its own copy until the process ends. In host code `scratch` names a host-side stand-in, not
// device memory: kernels get the address hipGetSymbolAddress returns for the device
// (g_scratch_dev), never the array's name.
constexpr int64_t kNoallocMaxKeys = (int64_t)1 << 22;  // 16 MiB of uint32
constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint32_t scratch[kNoallocMaxKeys];
static uint32_t* g_scratch_dev[kMaxDevices];  // the device address of `scratch`, per device

}  // namespace cilk

namespace host {

// Resolves the device address of the no-allocation workspace on `device`, which must be the
// current device. Called once per device by gtap_initialize.
int cilksort_scratch_setup(int device) {
    if (device < 0 || device >= cilk::kMaxDevices) return invalid_argument("cilksort_scratch_setup: device index");
    if (cilk::g_scratch_dev[device]) return 0;
    void* p = nullptr;
    GTAP_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(cilk::scratch)));
    if (p == nullptr) return fail(GTAP_ERROR_HIP, "hipGetSymbolAddress returned no address for the cilksort workspace");
    cilk::g_scratch_dev[device] = static_cast<uint32_t*>(p);
    return 0;
}

}  // namespace host
}  // namespace gtap

// gtap_cilksort with the library's static workspace instead of a caller's tmp buffer: no
// allocation of any kind, for callers that cannot allocate (under HIP graph capture, in a
// serving loop). The runtime must be initialized; n <= 2^22 keys. Calls on one GPU share the
// workspace, so calls that overlap on one GPU race.
extern "C" int gtap_cilksort_noalloc(uint32_t* keys, int64_t n, int64_t leaf, int64_t merge_leaf,
                                     gtap_run_stats* stats) {
    using namespace gtap;
    host::Runtime& R = host::rt();
    if (!R.initialized) return host::fail(GTAP_ERROR_NOT_INITIALIZED, "gtap_cilksort_noalloc before gtap_initialize");
    if (!keys || n < 0) return host::invalid_argument("gtap_cilksort_noalloc: bad arguments");
    if (n > cilk::kNoallocMaxKeys)
        return host::invalid_argument("gtap_cilksort_noalloc: n exceeds the workspace of 4194304 keys");
    uint32_t* tmp = (R.device >= 0 && R.device < cilk::kMaxDevices) ? cilk::g_scratch_dev[R.device] : nullptr;
    if (tmp == nullptr) return host::fail(GTAP_ERROR_NOT_INITIALIZED, "gtap_cilksort_noalloc: workspace not resolved");
    return gtap_cilksort(keys, tmp, n, leaf, merge_leaf, stats);
}

// Host-buffer convenience: copies keys in, sorts, copies the sorted keys back.
extern "C" int gtap_cilksort_host(uint32_t* keys, int64_t n, int64_t leaf, int64_t merge_leaf, gtap_run_stats* stats) {
    if (n < 0 || (n > 0 && !keys)) return gtap::host::invalid_argument("gtap_cilksort_host");
    if (n == 0) return gtap_cilksort(reinterpret_cast<uint32_t*>(16), reinterpret_cast<uint32_t*>(16), 0, leaf,
                                     merge_leaf, stats);
    uint32_t *dk = nullptr, *dt = nullptr;
    GTAP_HIP_TRY(hipMalloc(&dk, (size_t)n * 4));
    GTAP_HIP_TRY(hipMalloc(&dt, (size_t)n * 4));
    int st = 0;
    hipError_t e = hipMemcpy(dk, keys, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) st = gtap::host::hip_status(e, "gtap_cilksort_host copy-in");
    if (!st) st = gtap_cilksort(dk, dt, n, leaf, merge_leaf, stats);
    if (!st) {
        e = hipMemcpy(keys, dk, (size_t)n * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = gtap::host::hip_status(e, "gtap_cilksort_host copy-out");
    }
    (void)hipFree(dk);
    (void)hipFree(dt);
    return st;
}
