// gtap_thread.hpp -- thread-executed (per-lane) persistent scheduler for gfx950.
//
// Counterpart of the reference's thread mode (runtime/detail/thread/gtap_runtime_thread.cuh,
// gtap_thread_core.cuh), redesigned for wave64 and for CDNA4's per-XCD L2:
//
//  * Worker = one wavefront. A round runs up to 64 tasks, one per lane
//    (reference: 32 per warp, gtap_runtime_thread.cuh:849-1040).
//  * Task records are fixed-stride slots in HBM addressed by task id; a 16-byte header
//    (function id, resume state, EPAQ queue, parent id, join counter, result offset) is
//    followed by the frame. A child writes its return value straight into the slot the
//    parent named at spawn time, so nothing has to be retained after the join
//    (reference: retained child ids + frames, gtap_runtime_thread.cuh:749-754).
//  * Per worker and EPAQ queue there is one deque in HBM: a 64-bit word
//    (top << 32 | bottom) and a ring of task ids. The owner publishes with one atomic
//    add on the word and pops up to 64 of its newest tasks with one CAS; thieves take
//    up to half (<= 64) of the oldest with one CAS. One word arbitrates every claim,
//    so no lock is needed (reference: count + lock + head, gtap_runtime_thread.cuh:28-34,
//    511-593).
//  * Children are appended to the owner's ring *unpublished* while a round runs. After
//    the round the wave keeps the newest <= 64 of the busiest queue for its next round and
//    publishes the rest (reference: push_batch, gtap_runtime_thread.cuh:595-660).
//  * Task ids come from a wave-private stack in LDS, refilled and spilled in chunks of 64
//    through a bounded MPMC ring in HBM (reference: per-warp id lists,
//    gtap_thread_core.cuh:111-146).
//  * Joins are one agent-scope atomic decrement per finished child. The last child
//    appends the parent continuation to its own wave's queue given by the parent's
//    taskwait queue (reference: notify_parent, gtap_runtime_thread.cuh:714-730).
//  * Every cross-worker access is an sc1 (agent-coherent) access; there are no cache
//    fences (see gtap_dev.hpp).
#pragma once

#include <type_traits>

#include "gtap_dev.hpp"

namespace gtap {

constexpr uint32_t kFreeCap = 512;          // wave-private id stack (LDS)
constexpr uint32_t kRefillTarget = 192;     // refill the LDS stack up to this many ids
constexpr uint32_t kSpillTarget = 320;      // spill the LDS stack down to this many ids
constexpr uint32_t kPrivCap = 8192;         // wave-private overflow stack in HBM (ids)
constexpr uint32_t kChunk = 64;             // ids per pool cell
constexpr uint32_t kAnchorFunc = 0xFFFu;    // function id stored in anchor headers
constexpr uint32_t kMargin = 1024;          // deque slots kept free beyond the tail
constexpr uint32_t kIdleRegisterAfter = 4;  // failed acquire passes before a wave counts itself idle

enum Outcome : uint32_t { OUT_RUNNING = 0, OUT_FINISHED = 1, OUT_SUSPENDED = 2 };

// A program may name one EPAQ queue whose tasks run on a whole wavefront (P::kWaveQueue).
template <class P, class = void>
struct HasWaveQueue {
    static constexpr int value = -1;
};
template <class P>
struct HasWaveQueue<P, decltype((void)P::kWaveQueue)> {
    static constexpr int value = P::kWaveQueue;
};

// Per-wave scheduler state in LDS.
//
// The private end of each queue (positions [bottom, tail)) is mirrored in a ring of 64
// slots indexed by position & 63. A child spawned this round is held there *only in LDS*:
// its descriptor (header words and the first FW frame words) sits in desc[q][slot] and
// nothing is written to HBM. After the round the wave keeps the newest <= 64 entries of
// one queue for its next round and reads their descriptors back from LDS; only the
// entries it publishes, the ones pushed out of the 64-slot window by later appends, and
// the tasks that suspend at a taskwait are written to their HBM records. Most tasks of a
// fork-join tree never touch HBM for their spawn and start.
//   valid[q] bit s: slot s holds a private entry not yet published or taken;
//   inlds[q] bit s: that entry's record exists only as desc[q][s].
// The wave queue (run on a whole wave) keeps its records in HBM: QL = queues with
// descriptors. A descriptor is { word0, parent, result offset, frame[0..FW) } padded to an
// odd word count (LDS bank spread).
template <int Q, int QL, int D, int S, int NS>
struct WaveCtx {
    uint32_t free_top;
    uint32_t priv_top;
    uint32_t cur_queue;
    uint32_t spawns;
    uint32_t refills;
    uint32_t overflowed;    // LDS-held records written out by an append that pushed them out
    uint32_t ring_spilled;  // tasks moved to the overflow ring (deque near capacity)
    uint32_t tail[Q];       // owner's private tail per queue (absolute position)
    uint32_t bottom[Q];     // published bottom per queue (low word of the deque word)
    uint32_t top_cache[Q];  // last top observed per queue
    uint64_t valid[Q];
    uint64_t inlds[Q];
    uint32_t free_ids[kFreeCap];
    uint32_t mirror[Q][kWave];  // ids of the last 64 entries appended per queue, by position & 63
    uint32_t desc[QL][kWave][D];
    // records written out together after a round (suspended tasks, published entries): the
    // 16-byte header and S/4 - 1 frame pieces each, spread over the lanes so that the pieces
    // of one record leave in one store instruction and merge into one memory transaction
    uint32_t stage[NS][S];
    uint32_t stage_id[NS];
    uint32_t nstage;        // records staged so far this round (entries pushed out of the window, <= kEvictStage)
};

// A program whose tasks never wait for children (P::kNoTaskwait) never suspends a task, so
// a round stages at most the 64 records pushed out of the window: half the staging area.
template <class P, class = void>
struct NoTaskwait {
    static constexpr bool value = false;
};
template <class P>
struct NoTaskwait<P, decltype((void)P::kNoTaskwait)> {
    static constexpr bool value = P::kNoTaskwait;
};

// Staging entries per wave (evicted records, plus a round's suspended tasks for programs that
// wait for children): two rounds' worth by default; a program short of LDS may ask for 64
// (P::kStageEntries), and its evicted records are then written out at once instead of staged.
template <class P, class = void>
struct StageEntries {
    static constexpr int value = NoTaskwait<P>::value ? kWave : 2 * kWave;
};
template <class P>
struct StageEntries<P, decltype((void)P::kStageEntries)> {
    static constexpr int value = P::kStageEntries;
};
template <class P>
struct CtxShape {
    static constexpr int kQ = P::kQueues;
    static constexpr int kWQ = HasWaveQueue<P>::value;
    static_assert(kWQ < 0 || kWQ == kQ - 1, "the wave queue must be the last EPAQ queue");
    static constexpr int kQL = kWQ >= 0 ? kQ - 1 : kQ;  // queues whose entries may live in LDS
    static constexpr int kFW = P::kFrameWords;
    static constexpr int kFWr = (kFW + 3) & ~3;     // frame words kept in registers (16-byte pieces)
    static constexpr int kPieces = 1 + kFWr / 4;    // 16-byte pieces of header + frame words
    static_assert(kFWr <= 12, "at most 12 frame words are loaded with a task (one 64-byte line)");
    static_assert(P::kFrameBytes >= 4 * kFWr, "the frame must cover the loaded frame words");
    static constexpr int kD = (3 + kFW) | 1;
    static constexpr int kStage = StageEntries<P>::value;  // evicted (+ suspended) records
    static_assert(kStage >= kWave, "the staging area holds at least one round's suspended tasks");
    // evicted records staged per round; the rest of the area is for the suspended tasks
    static constexpr uint32_t kEvictStage = (uint32_t)kStage - (NoTaskwait<P>::value ? 0u : (uint32_t)kWave);
    using type = WaveCtx<kQ, (kQL > 0 ? kQL : 1), kD, 4 * kPieces, kStage>;
};
template <class P>
using WaveCtxOf = typename CtxShape<P>::type;

__device__ __forceinline__ uint8_t* rec_ptr(const GtapDev& dev, uint32_t id) {
    return dev.recs + (size_t)id * dev.stride;
}
__device__ __forceinline__ uint64_t* dq_word_ptr(const GtapDev& dev, uint32_t q, uint32_t w) {
    return dev.dq_word + ((size_t)q * dev.nworkers + w) * 16;
}
__device__ __forceinline__ uint32_t* dq_slot_ptr(const GtapDev& dev, uint32_t q, uint32_t w, uint32_t pos) {
    return dev.dq_slots + ((size_t)q * dev.nworkers + w) * dev.cap + (pos & (dev.cap - 1));
}

// ------------------------------------------------------------------ id pool
// Three levels. (1) A 512-id stack per wave in LDS. (2) An 8192-id overflow stack per
// wave in HBM, touched only by its owner (no atomics). (3) `nshards` bounded MPMC rings of
// 64-id cells (Vyukov's sequence numbers) that move ids between waves, plus a counter of
// never-used ids. Every ring holds all ids, so an enqueue ticket (fetch-add) never has to
// wait for more than a dequeuer that is still copying the cell out.
// Ring s: pool_ctl[32s] = enqueue position, pool_ctl[32s+16] = dequeue position;
// pool_ctl[32*nshards] = fresh-id counter. Cell c of lap L is full when seq == pos+1.

__device__ __forceinline__ uint32_t* pool_enq_ptr(const GtapDev& dev, uint32_t s) {
    return reinterpret_cast<uint32_t*>(dev.pool_ctl + (size_t)s * 32);
}
__device__ __forceinline__ uint32_t* pool_deq_ptr(const GtapDev& dev, uint32_t s) {
    return reinterpret_cast<uint32_t*>(dev.pool_ctl + (size_t)s * 32 + 16);
}
__device__ __forceinline__ uint32_t* pool_fresh_ptr(const GtapDev& dev) {
    return reinterpret_cast<uint32_t*>(dev.pool_ctl + (size_t)dev.nshards * 32);
}
__device__ __forceinline__ uint32_t* pool_seq_ptr(const GtapDev& dev, uint32_t s, uint32_t pos) {
    return dev.pool_seq + (size_t)s * dev.ncells + (pos & (dev.ncells - 1));
}
__device__ __forceinline__ uint32_t* pool_cell_ptr(const GtapDev& dev, uint32_t s, uint32_t pos) {
    return dev.pool_ids + ((size_t)s * dev.ncells + (pos & (dev.ncells - 1))) * kChunk;
}

// Leader lane only: claim a full cell of ring s (a few CAS attempts; false when the ring
// is empty or contended).
__device__ __forceinline__ bool ring_take_claim(const GtapDev& dev, uint32_t s, uint32_t& pos_out) {
    uint32_t* deq = pool_deq_ptr(dev, s);
    for (int tries = 0; tries < 4; ++tries) {
        const uint32_t pos = ld32(deq);
        const uint32_t seq = ld32(pool_seq_ptr(dev, s, pos));
        const int32_t diff = (int32_t)(seq - (pos + 1));
        if (diff == 0) {
            uint32_t expected = pos;
            if (atomic_cas(deq, expected, pos + 1)) {
                pos_out = pos;
                return true;
            }
        } else if (diff < 0) {
            return false;
        }
    }
    return false;
}

// Leader lane only: take an enqueue ticket of ring s and wait until its cell is free.
__device__ __forceinline__ bool ring_put_claim(const GtapDev& dev, uint32_t s, uint32_t& pos_out) {
    const uint32_t pos = atomic_add(pool_enq_ptr(dev, s), 1u);
    for (uint32_t spin = 0; spin < (1u << 22); ++spin) {
        if (ld32(pool_seq_ptr(dev, s, pos)) == pos) {
            pos_out = pos;
            return true;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return false;
}

// Active lanes look at the rings together, lane of rank r at ring (me + b + r) % nshards
// for b = 0, cnt, 2 cnt, ...: its dequeue position and that cell's sequence number. Returns
// (wave-uniform) the first ring seen with a full cell at its dequeue position.
__device__ __forceinline__ bool ring_probe(const GtapDev& dev, uint32_t me, uint32_t cnt, uint32_t r,
                                           uint32_t& shard_out, uint32_t& pos_out) {
    for (uint32_t b = 0; b < dev.nshards; b += cnt) {
        const uint32_t i = b + r;
        bool full = false;
        uint32_t sh = 0, pos = 0;
        if (i < dev.nshards) {
            sh = (me + i) % dev.nshards;
            pos = ld32(pool_deq_ptr(dev, sh));
            full = ld32(pool_seq_ptr(dev, sh, pos)) == pos + 1;
        }
        const uint64_t f = ballot(full);
        if (f) {
            const int src = first_lane(f);
            shard_out = shfl(sh, src);
            pos_out = shfl(pos, src);
            return true;
        }
    }
    return false;
}

// Add ids to the wave's LDS stack: one chunk from the wave's HBM stack, else one ring cell,
// else `want` chunks (fewer when the pool runs out) of never-used ids with one fetch-add.
// Called by all active lanes.
template <class W>
__device__ __forceinline__ void refill_chunk(const GtapDev& dev, W* w, uint32_t me, uint32_t want) {
    const uint64_t m = active_mask();
    const uint32_t cnt = (uint32_t)popc(m);
    const uint32_t r = (uint32_t)rank_in(m);
    const int leader = first_lane(m);
    const uint32_t top = uniform(vld(&w->free_top));
    const uint32_t ptop = uniform(vld(&w->priv_top));
    if (ptop >= kChunk) {
        const uint32_t* src = dev.priv_ids + (size_t)me * kPrivCap + (ptop - kChunk);
        for (uint32_t i = r; i < kChunk; i += cnt) vst(&w->free_ids[top + i], src[i]);
        vst(&w->priv_top, ptop - kChunk);
        vst(&w->free_top, top + kChunk);
        return;
    }
    uint32_t pos = 0, shard = 0, fresh = 0, nch = 0;
    int kind = 0;  // 0 = exhausted, 1 = ring cell, 2 = fresh ids
    bool fresh_left = true;
    for (int attempt = 0; attempt < 64; ++attempt) {
        const bool seen = ring_probe(dev, me, cnt, r, shard, pos);
        if (seen) {
            int ok = 0;
            if ((int)r == 0) {
                uint32_t expected = pos;
                ok = atomic_cas(pool_deq_ptr(dev, shard), expected, pos + 1) ? 1 : 0;
            }
            if (shfl((int32_t)ok, leader)) {
                kind = 1;
                nch = 1;
                break;
            }
        }
        if (fresh_left) {
            uint32_t f = 0, got = 0;
            if ((int)r == 0) {
                f = atomic_add(pool_fresh_ptr(dev), want * kChunk);
                const uint32_t left = f < dev.nrecs ? (dev.nrecs - f) / kChunk : 0u;
                got = left < want ? left : want;
            }
            fresh = shfl(f, leader);
            nch = shfl(got, leader);
            if (nch) {
                kind = 2;
                break;
            }
            fresh_left = false;
        }
        if (!seen) break;  // every ring empty and no never-used ids left
    }
    if (kind == 1) {
        const uint32_t* ids = pool_cell_ptr(dev, shard, pos);
        for (uint32_t i = r; i < kChunk; i += cnt) vst(&w->free_ids[top + i], ld32(ids + i));
        complete_stores();  // the cell is read out before it is released
        if ((int)r == 0) st32(pool_seq_ptr(dev, shard, pos), pos + dev.ncells);
    } else if (kind == 2) {
        for (uint32_t i = r; i < nch * kChunk; i += cnt) vst(&w->free_ids[top + i], fresh + i);
    } else {
        if ((int)r == 0) GTAP_DEVICE_ERROR(dev, GTAP_ERROR_TASK_ID_POOL_LOW_HEADROOM, -1, -1, (int)fresh, (int)dev.nrecs);
        // hand out the sink record (id 0) so the round can finish harmlessly
        nch = 1;
        for (uint32_t i = r; i < kChunk; i += cnt) vst(&w->free_ids[top + i], 0u);
    }
    if ((int)r == 0) vst(&w->refills, vld(&w->refills) + 1u);
    vst(&w->free_top, top + nch * kChunk);
}

// Move the 64 most recently freed ids out of the wave's LDS stack.
template <class W>
__device__ __forceinline__ void spill_chunk(const GtapDev& dev, W* w, uint32_t me) {
    const uint64_t m = active_mask();
    const uint32_t cnt = (uint32_t)popc(m);
    const uint32_t r = (uint32_t)rank_in(m);
    const int leader = first_lane(m);
    const uint32_t top = uniform(vld(&w->free_top));
    const uint32_t ptop = uniform(vld(&w->priv_top));
    if (ptop + kChunk <= dev.priv_cap) {
        uint32_t* dst = dev.priv_ids + (size_t)me * kPrivCap + ptop;
        for (uint32_t i = r; i < kChunk; i += cnt) dst[i] = vld(&w->free_ids[top - kChunk + i]);
        vst(&w->priv_top, ptop + kChunk);
    } else {
        const uint32_t shard = me % dev.nshards;
        uint32_t pos = 0;
        int ok = 0;
        if ((int)r == 0) ok = ring_put_claim(dev, shard, pos) ? 1 : 0;
        ok = shfl((int32_t)ok, leader);
        pos = shfl(pos, leader);
        if (ok) {
            uint32_t* ids = pool_cell_ptr(dev, shard, pos);
            for (uint32_t i = r; i < kChunk; i += cnt) st32(ids + i, vld(&w->free_ids[top - kChunk + i]));
            complete_stores();
            if ((int)r == 0) st32(pool_seq_ptr(dev, shard, pos), pos + 1);
        } else if ((int)r == 0) {
            GTAP_DEVICE_ERROR(dev, GTAP_ERROR_TASK_ID_POOL_SLOT_BUSY, -1, (int)shard, (int)pos, (int)dev.ncells);
        }
    }
    vst(&w->free_top, top - kChunk);
}

template <class W>
__device__ __forceinline__ uint32_t alloc_id(const GtapDev& dev, W* w, uint32_t me) {
    const uint64_t m = active_mask();
    const uint32_t cnt = (uint32_t)popc(m);
    const uint32_t r = (uint32_t)rank_in(m);
    uint32_t top = uniform(vld(&w->free_top));
    if (top < cnt) {
        while (top < kRefillTarget) {
            refill_chunk(dev, w, me, (kRefillTarget - top + kChunk - 1) / kChunk);
            top = uniform(vld(&w->free_top));
        }
    }
    const uint32_t id = vld(&w->free_ids[top - 1 - r]);
    vst(&w->free_top, top - cnt);
    return id;
}

template <class W>
__device__ __forceinline__ void free_id(const GtapDev& dev, W* w, uint32_t me, uint32_t id) {
    const uint64_t m = active_mask();
    const uint32_t cnt = (uint32_t)popc(m);
    const uint32_t r = (uint32_t)rank_in(m);
    uint32_t top = uniform(vld(&w->free_top));
    if (top + cnt > kFreeCap) {
        while (top > kSpillTarget) {
            spill_chunk(dev, w, me);
            top = uniform(vld(&w->free_top));
        }
    }
    vst(&w->free_ids[top + r], id);
    vst(&w->free_top, top + cnt);
}

// ------------------------------------------------------------------ deque append

__device__ __forceinline__ uint64_t rotl64(uint64_t x, uint32_t r) {
    r &= 63u;
    return r ? ((x << r) | (x >> (64u - r))) : x;
}
__device__ __forceinline__ uint64_t low_bits(uint32_t n) { return n >= 64u ? ~0ull : ((1ull << n) - 1ull); }

// Writes the HBM record of an entry held in LDS: the 16-byte header (join counter 0) and the
// first FW frame words.
template <class P>
__device__ __forceinline__ void record_from_desc(const GtapDev& dev, __amdgpu_buffer_rsrc_t rs, uint32_t id,
                                                 const uint32_t* d) {
    constexpr int FW = CtxShape<P>::kFW;
    const uint32_t off = id * dev.stride;
    uint32_t f[FW];
#pragma unroll
    for (int i = 0; i < FW; ++i) f[i] = vld(&d[3 + i]);
    st128(rs, off, u32x4{vld(&d[1]), vld(&d[2]), 0u, vld(&d[0])});
    frame_words_st<FW>(rs, off, f);
}

// Copies an LDS-held entry into staging entry `at` in record order (header, frame words).
template <class P, class W>
__device__ __forceinline__ void stage_from_desc(W* w, uint32_t at, uint32_t id, const uint32_t* d) {
    constexpr int FW = CtxShape<P>::kFW, FWr = CtxShape<P>::kFWr;
    uint32_t* dst = &w->stage[at][0];
    vst(&dst[0], vld(&d[1]));
    vst(&dst[1], vld(&d[2]));
    vst(&dst[2], 0u);
    vst(&dst[3], vld(&d[0]));
#pragma unroll
    for (int i = 0; i < FWr; ++i) vst(&dst[4 + i], (i < FW) ? vld(&d[3 + (i < FW ? i : 0)]) : 0u);
    vst(&w->stage_id[at], id);
}

// ------------------------------------------------------------------ deque overflow ring
// One MPMC ring of task ids per EPAQ queue takes published tasks off a deque that nears its
// capacity, where the reference traps (runtime/detail/thread/gtap_runtime_thread.cuh:42-45).
// A cell is one 64-bit word: state [62..63] | position tag [32..61] (position mod 2^30) |
// task id [0..31]. The cell of position p is free for p when it reads FREE with tag p. A
// producer takes positions with a fetch-add on the ticket counter and, once the task's
// record is complete, writes READY | p | id; a consumer claims a run of READY cells at the
// head with one CAS, reads the ids and frees each cell for p + cap. Positions only grow:
// the counters are never reset, and the reset kernel frees what an early end leaves between
// the head and the tickets. Idle waves take from the rings. Every push is followed by a bump
// of 64 epoch words, one per group of waves (wave me reads word me % 64, so the reads spread
// over 64 lines): a wave looks at the rings only when its group's epoch moved since it last
// found them all empty, and counts itself idle only when it has not moved. A wave that pushed
// therefore sees its own push before it can count itself idle, and quiescence still implies
// that no task is anywhere.
constexpr uint64_t kOvfReady = 2ull << 62;
constexpr uint64_t kOvfTagMask = 0x3FFFFFFFull;
constexpr uint32_t kOvfSpillMax = 4096;  // tasks one deque spill moves at most
__device__ __forceinline__ uint64_t ovf_free_word(uint64_t pos) { return (pos & kOvfTagMask) << 32; }
__device__ __forceinline__ uint64_t ovf_ready_word(uint64_t pos, uint32_t id) {
    return kOvfReady | ((pos & kOvfTagMask) << 32) | (uint64_t)id;
}

// What the ring code uses of the launch, by value: the ring code is out of line (none of it
// sits in the hot loop's registers) and takes scalars only, so the kernel argument block is
// never copied to scratch for it.
struct OvfView {
    uint64_t* cells;
    uint64_t* ctl;
    int32_t* dctl;
    gtap_error_report* report;
    uint32_t* epoch;
    uint32_t cap;
    uint32_t term_mode;
};
__device__ __forceinline__ OvfView ovf_view(const GtapDev& dev) {
    return OvfView{dev.ovf_cells, dev.ovf_ctl, dev.ctl, dev.report, dev.ovf_epoch, dev.ovf_cap, dev.term_mode};
}
constexpr uint32_t kOvfEmpty = 0x40000000u;  // ovf_pop: the ring held no tickets
constexpr uint32_t kOvfLeftIdle = 0x80000000u;  // ovf_pop: the wave left the idle count
__device__ __forceinline__ uint64_t* ovf_cell(const OvfView& v, uint32_t q, uint64_t pos) {
    return v.cells + (size_t)q * v.cap + (pos & (uint64_t)(v.cap - 1));
}
__device__ __forceinline__ uint64_t* ovf_head(const OvfView& v, uint32_t q) { return v.ctl + (size_t)q * 32; }
__device__ __forceinline__ uint64_t* ovf_tickets(const OvfView& v, uint32_t q) { return v.ctl + (size_t)q * 32 + 16; }

// Positions [base, base + n) of ring q were just taken: false (after recording the error) if
// they would come within kMargin cells of the head's lap.
__device__ __forceinline__ bool ovf_room(const OvfView& v, uint32_t q, uint64_t base, uint32_t n) {
    const uint64_t head = ld64(ovf_head(v, q));
    if (base + n - head > (uint64_t)(v.cap - kMargin)) {
        record_error(v.dctl, v.report, GTAP_ERROR_QUEUE_OVERFLOW, -1, (int)q, (int)(base + n - head),
                     (int)(v.cap - kMargin), __LINE__);
        return false;
    }
    return true;
}

// Writes task `id` (record complete) into cell `pos` of ring q. The position's previous lap
// was claimed before ovf_room let the position out, but its consumer may still be reading
// the cell: wait until it is free (the consumer frees it right after its read).
__device__ __forceinline__ void ovf_fill(const OvfView& v, uint32_t q, uint64_t pos, uint32_t id) {
    uint64_t* c = ovf_cell(v, q, pos);
    const uint64_t want = ovf_free_word(pos);
    for (uint32_t spin = 0; ld64(c) != want; ++spin) {
        if ((spin & 255u) == 255u && ld32(v.dctl + CTL_ABORT * 32)) return;
        if (spin > (1u << 22)) {
            record_error(v.dctl, v.report, GTAP_ERROR_WATCHDOG, (int)id, (int)q, (int)(pos & 0x7FFFFFFF), 0, __LINE__);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    st64(c, ovf_ready_word(pos, id));
}

// Moves the oldest published tasks of a deque to ring q: called by the whole wave after its
// round's publish when the deque holds more than half its capacity. One CAS on the deque word
// claims them from the top end, as a thief would, then the wave writes their ids into ring
// cells. `word` / `slots` are the deque's word and slot array, `bottom` its published bottom.
// Returns the top the deque word was seen at (or moved to) and, in the high half, the number
// of tasks moved.
__device__ __forceinline__ uint64_t ovf_spill_deque(OvfView v, uint32_t q, uint64_t* word, uint32_t* slots,
                                                 uint32_t dq_cap) {
    const int lane = lane_id();
    const uint32_t half = dq_cap / 2;
    uint64_t wv = 0;
    if (lane == 0) wv = ld64(word);
    wv = shfl(wv, 0);
    const uint32_t top = (uint32_t)(wv >> 32), bw = (uint32_t)wv;
    if ((int32_t)(bw - top) <= (int32_t)half) return top;
    uint32_t k = (bw - top) - half / 2;
    if (k > kOvfSpillMax) k = kOvfSpillMax;
    int32_t ok = 0;
    uint64_t base = 0;
    if (lane == 0) {
        uint64_t expected = wv;
        ok = atomic_cas(word, expected, wv + ((uint64_t)k << 32)) ? 1 : 0;
        if (ok) {
            base = atomic_add(ovf_tickets(v, q), (uint64_t)k);
            if (!ovf_room(v, q, base, k)) ok = 2;
        }
    }
    ok = shfl(ok, 0);
    if (ok == 0) return top;                 // a thief moved the top: the next round tries again
    if (ok == 2) return top + k;             // the ring is full: the error ends the launch
    base = shfl(base, 0);
    for (uint32_t j = (uint32_t)lane; j < k; j += kWave) ovf_fill(v, q, base + j, ld32(slots + ((top + j) & (dq_cap - 1))));
    complete_stores();
    atomic_add_noret(v.epoch + (uint32_t)lane * 32u, 1u);  // every group's epoch (one lane each)
    return (uint64_t)(top + k) | ((uint64_t)k << 32);
}

// Takes up to `cap` ready tasks (64, or 1 on the wave queue) from the head of ring q. Low
// half of the result: the number taken (lane i < that number receives one in the high half),
// or kOvfEmpty when the ring held no tickets; kOvfLeftIdle: the wave left the idle count
// before its claim.
__device__ __forceinline__ uint64_t ovf_pop(OvfView v, uint32_t q, uint32_t cap, int32_t idle_registered) {
    const int lane = lane_id();
    uint64_t h = 0, tk = 0;
    if (lane == 0) {
        h = ld64(ovf_head(v, q));
        tk = ld64(ovf_tickets(v, q));
    }
    h = shfl(h, 0);
    tk = shfl(tk, 0);
    if ((int64_t)(tk - h) <= 0) return kOvfEmpty;
    const uint32_t avail = (tk - h) < (uint64_t)cap ? (uint32_t)(tk - h) : cap;
    uint64_t cv = 0;
    bool ready = false;
    if ((uint32_t)lane < avail) {
        cv = ld64(ovf_cell(v, q, h + (uint32_t)lane));
        ready = cv == ovf_ready_word(h + (uint32_t)lane, (uint32_t)cv);
    }
    const uint64_t notready = ~ballot(ready);
    const uint32_t k = notready ? (uint32_t)__builtin_ctzll(notready) : 64u;  // ready run from lane 0
    if (k == 0) return 0;
    uint32_t left = 0;
    if (v.term_mode == GTAP_TERM_QUIESCE && idle_registered) {
        if (lane == 0) atomic_add(v.dctl + CTL_IDLE * 32, -1);
        left = kOvfLeftIdle;
    }
    int32_t ok = 0;
    if (lane == 0) {
        uint64_t expected = h;
        ok = atomic_cas(ovf_head(v, q), expected, h + k) ? 1 : 0;
    }
    ok = shfl(ok, 0);
    if (!ok) return left;
    uint32_t tid = 0;
    if ((uint32_t)lane < k) {
        tid = (uint32_t)cv;
        st64(ovf_cell(v, q, h + (uint32_t)lane), ovf_free_word(h + (uint32_t)lane + v.cap));
    }
    return (uint64_t)(k | left) | ((uint64_t)tid << 32);
}

// Append `id` to the private (unpublished) end of queue `q` of this wave. Called by all
// active lanes; lanes may target different queues. kDesc: the entry's record is held in
// LDS (the D descriptor words `dw`), else it is already in HBM. An older private entry that
// the append pushes out of the 64-slot window is written out: its record if it was held in
// LDS, and its deque slot. A round that appends more than the deque holds beyond its
// published tasks ends the launch with GTAP_ERROR_QUEUE_OVERFLOW (the round's children are
// private until the round ends: their parents' taskwait states are written then); what
// accumulates over rounds goes to the overflow ring instead (ovf_spill_deque).
template <class P, bool kDesc>
__device__ __forceinline__ void queue_append(const GtapDev& dev, WaveCtxOf<P>* w, uint32_t me, uint32_t q,
                                             uint32_t id, const uint32_t* dw) {
    constexpr int Q = CtxShape<P>::kQ, QL = CtxShape<P>::kQL, D = CtxShape<P>::kD;
    const __amdgpu_buffer_rsrc_t rs = pool_rsrc(dev.recs);
#pragma unroll
    for (int qq = 0; qq < Q; ++qq) {
        const uint64_t mq = ballot(q == (uint32_t)qq);
        if (mq == 0) continue;
        if (q == (uint32_t)qq) {
            const uint32_t cnt = (uint32_t)popc(mq);
            const uint32_t r = (uint32_t)rank_in(mq);
            const uint32_t t = uniform(vld(&w->tail[qq]));
            uint32_t top = uniform(vld(&w->top_cache[qq]));
            if (t + cnt - top > dev.cap - kMargin) {
                // refresh the cached top from the deque word before declaring overflow
                uint64_t wv = ld64(dq_word_ptr(dev, qq, me));
                top = uniform((uint32_t)(wv >> 32));
                vst(&w->top_cache[qq], top);
            }
            if (t + cnt - top > dev.cap - kMargin) {
                if (r == 0) GTAP_DEVICE_ERROR(dev, GTAP_ERROR_QUEUE_OVERFLOW, id, qq, t + cnt - top, dev.cap - kMargin);
            } else {
                const uint32_t pos = t + r;
                const uint32_t slot = pos & (kWave - 1);
                const uint64_t vmask = uniform(vld(&w->valid[qq]));
                const uint64_t lmask = uniform(vld(&w->inlds[qq]));
                const bool spill = (vmask >> slot) & 1ull;
                const bool spill_rec = spill && qq < QL && ((lmask >> slot) & 1ull);
                const uint64_t om = ballot(spill_rec);
                if (spill) {
                    // the slot still holds the private entry at pos - 64: its deque slot now,
                    // its record (if held in LDS) staged for the writes after the round, or
                    // written at once when the staging area is full
                    const uint32_t oid = vld(&w->mirror[qq][slot]);
                    if (spill_rec) {
                        const uint32_t base = uniform(vld(&w->nstage));
                        const uint32_t at = base + (uint32_t)rank_in(om);
                        const uint32_t* d = &w->desc[qq < QL ? qq : 0][slot][0];
                        if (base + (uint32_t)popc(om) <= CtxShape<P>::kEvictStage) {  // + <= 64 suspended after the round
                            stage_from_desc<P>(w, at, oid, d);
                        } else {
                            record_from_desc<P>(dev, rs, oid, d);
                        }
                    }
                    st32(dq_slot_ptr(dev, qq, me, pos - kWave), oid);
                }
                if (om && r == 0) {
                    const uint32_t base = vld(&w->nstage);
                    if (base + (uint32_t)popc(om) <= CtxShape<P>::kEvictStage) vst(&w->nstage, base + (uint32_t)popc(om));
                    vst(&w->overflowed, vld(&w->overflowed) + (uint32_t)popc(om));
                }
                vst(&w->mirror[qq][slot], id);
                if constexpr (kDesc) {
                    if (qq < QL) {
#pragma unroll
                        for (int i = 0; i < D; ++i) vst(&w->desc[qq < QL ? qq : 0][slot][i], dw[i]);
                    }
                }
                const uint64_t range = rotl64(low_bits(cnt), t);
                if (r == 0) {
                    vst(&w->valid[qq], vmask | range);
                    vst(&w->inlds[qq], (kDesc && qq < QL) ? (lmask | range) : (lmask & ~range));
                    vst(&w->tail[qq], t + cnt);
                }
            }
        }
    }
}

// ------------------------------------------------------------------ spawn

// Allocate a task id and append the child to the private end of queue q: held in LDS (a
// descriptor) on a per-lane queue, written to its HBM record on the wave queue. Called by
// every lane that spawns (wave-aggregated inside).
template <class P, typename A>
__device__ __forceinline__ uint32_t spawn_record(const GtapDev& dev, WaveCtxOf<P>* w, uint32_t me,
                                                 uint32_t self, uint32_t f, uint32_t q, uint32_t par, uint32_t roff,
                                                 const A& a) {
    static_assert(sizeof(A) % 8 == 0, "argument blocks are padded to 8 bytes");
    static_assert(sizeof(A) / 4 <= P::kFrameWords, "arguments must fit the loaded frame");
    constexpr int D = CtxShape<P>::kD, WQ = CtxShape<P>::kWQ, FW = CtxShape<P>::kFW;
    if (q >= (uint32_t)P::kQueues) {
        GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_QUEUE_IDX, self, q, q, P::kQueues);
        q = 0;
    }
    const uint32_t id = alloc_id(dev, w, me);
    uint32_t d[D];
    d[0] = hdr_word0(f, 0, q);
    d[1] = par;
    d[2] = roff;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&a);
#pragma unroll
    for (int i = 0; i < FW; ++i) d[3 + i] = (i < (int)(sizeof(A) / 4)) ? src[i < (int)(sizeof(A) / 4) ? i : 0] : 0u;
#pragma unroll
    for (int i = 3 + FW; i < D; ++i) d[i] = 0u;
    if (WQ >= 0 && q == (uint32_t)WQ) {
        const __amdgpu_buffer_rsrc_t rs = pool_rsrc(dev.recs);
        const uint32_t off = id * dev.stride;
        st128(rs, off, u32x4{d[1], d[2], 0u, d[0]});
        frame_words_st<FW>(rs, off, &d[3]);
        queue_append<P, false>(dev, w, me, q, id, nullptr);
    } else {
        queue_append<P, true>(dev, w, me, q, id, d);
    }
    const uint64_t am = active_mask();
    if (rank_in(am) == 0) vst(&w->spawns, vld(&w->spawns) + (uint32_t)popc(am));
    return id;
}

// Memory ordering of LDS between the lanes of one wave (wave-cooperative tasks).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Writes the `count` records staged in w->stage / w->stage_id: piece k (16 bytes) of entry e
// goes out from lane e * kPieces + k, so one store instruction carries 64 / kPieces whole
// records and the pieces of a record merge into one memory transaction. An entry whose id
// carries kHeaderOnly writes its header piece only. All lanes of the wave call it.
constexpr uint32_t kHeaderOnly = 0x80000000u;
template <class P, class W>
__device__ __forceinline__ void write_staged(const GtapDev& dev, __amdgpu_buffer_rsrc_t rs, W* w, uint32_t count) {
    constexpr uint32_t PC = CtxShape<P>::kPieces;
    constexpr uint32_t E = kWave / PC;  // records per store instruction
    const uint32_t lane = (uint32_t)lane_id();
    const uint32_t k = lane % PC;
    for (uint32_t base = 0; base < count; base += E) {
        const uint32_t e = base + lane / PC;
        if (lane < E * PC && e < count) {
            const uint32_t v = w->stage_id[e];
            if (k == 0 || !(v & kHeaderOnly)) {
                const uint32_t* src = &w->stage[e][4 * k];
                st128(rs, (v & ~kHeaderOnly) * dev.stride + 16u * k, u32x4{src[0], src[1], src[2], src[3]});
            }
        }
    }
}

// ------------------------------------------------------------------ per-lane task context

template <class P>
struct Lane {
    static constexpr int Q = P::kQueues;
    static constexpr int FW = CtxShape<P>::kFWr;
    const GtapDev* dev;
    WaveCtxOf<P>* w;
    const typename P::Args* args;
    __amdgpu_buffer_rsrc_t rs;  // the record pool
    uint8_t* rec;           // this task's record
    uint32_t me;            // worker (wave) index
    uint32_t tid;
    uint32_t func;
    uint32_t state;
    uint32_t queue;
    uint32_t parent;
    uint32_t res_off;
    uint32_t child_count;
    uint32_t child_sum;     // of which children of an additive join
    uint32_t join_count;    // children to wait for, after a taskwait that suspends
    uint32_t outcome;
    bool lds;               // the record's header and first FW frame words are not in HBM yet
    bool sum_join;          // the pending taskwait is an additive join
    int64_t jsum;           // sum of the children's results, after an additive join
    uint64_t join_delta;    // what finishing adds to the parent's join word
    uint32_t frame[FW];     // the first frame words, loaded when the round starts

    // -- reads of the loaded frame (arguments written by the spawner)
    template <typename T>
    __device__ __forceinline__ T arg(int byte_off) const {
        static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4/8-byte args");
        if constexpr (sizeof(T) == 4) {
            return __builtin_bit_cast(T, frame[byte_off / 4]);
        } else {
            uint64_t v = (uint64_t)frame[byte_off / 4] | ((uint64_t)frame[byte_off / 4 + 1] << 32);
            return __builtin_bit_cast(T, v);
        }
    }
    // A task started from LDS writes its first FW frame words to HBM before its frame is
    // accessed there (the header follows at taskwait).
    __device__ __forceinline__ void materialize_frame() {
        frame_words_st<FW>(rs, tid * dev->stride, frame);
        lds = false;
    }
    // The 16-byte header this task leaves in its record when it suspends.
    __device__ __forceinline__ u32x4 suspended_header() const {
        if (sum_join)
            return u32x4{parent, (res_off & 0xFFFFu) | (state << 16) | (func << 24) | kSumLayout, join_count | (queue << 12), 0u};
        return u32x4{parent, res_off, join_count, hdr_word0(func, state, queue)};
    }
    // -- coherent access to any part of this task's frame (offsets relative to the frame)
    template <typename T>
    __device__ __forceinline__ T frame_ld(int byte_off) {
        if (lds) {
            materialize_frame();
            complete_stores();
        }
        return ldT<T>(rec + kHdrBytes + byte_off);
    }
    // A store to a field of the loaded frame words (constant offset) also updates the
    // register copy, so the task reads it back with arg() when it resumes; an LDS-held task
    // writes it to HBM with the rest of its frame words at taskwait.
    template <typename T>
    __device__ __forceinline__ void frame_st(int byte_off, T v) {
        static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4/8-byte fields");
        if (__builtin_constant_p(byte_off) && byte_off % 4 == 0 && byte_off + (int)sizeof(T) <= FW * 4) {
            if constexpr (sizeof(T) == 4) {
                frame[byte_off / 4] = __builtin_bit_cast(uint32_t, v);
            } else {
                const uint64_t u = __builtin_bit_cast(uint64_t, v);
                frame[byte_off / 4] = (uint32_t)u;
                frame[byte_off / 4 + 1] = (uint32_t)(u >> 32);
            }
            if (!lds) stT<T>(rec + kHdrBytes + byte_off, v);
            return;
        }
        if (lds) materialize_frame();
        stT<T>(rec + kHdrBytes + byte_off, v);
    }
    // Two adjacent 8-byte fields at a 16-byte aligned frame offset, with one 16-byte load.
    template <typename T>
    __device__ __forceinline__ void frame_ld2(int byte_off, T& a, T& b) {
        static_assert(sizeof(T) == 8, "8-byte fields");
        if (lds) {
            materialize_frame();
            complete_stores();
        }
        const u32x4 v = ld128(rs, tid * dev->stride + kHdrBytes + (uint32_t)byte_off);
        a = __builtin_bit_cast(T, (uint64_t)v.x | ((uint64_t)v.y << 32));
        b = __builtin_bit_cast(T, (uint64_t)v.z | ((uint64_t)v.w << 32));
    }

    __device__ __forceinline__ uint32_t get_state() const { return state; }

    // Spawn a child. `res_off` is the byte offset (inside this task's frame) of the slot
    // that receives the child's return value. Returns the child id.
    template <typename A>
    __device__ __forceinline__ uint32_t spawn(uint32_t f, uint32_t q, int res_off_in_frame, const A& a) {
        return spawn_impl(f, q, tid, kHdrBytes + (uint32_t)res_off_in_frame, a);
    }
    // Spawn a child nobody joins (the parent may return without taskwait).
    template <typename A>
    __device__ __forceinline__ uint32_t spawn_detached(uint32_t f, uint32_t q, const A& a) {
        return spawn_impl(f, q, kNoParent, 0u, a);
    }
    // Spawn a child whose integer result is added into this task's join sum (taskwait_sum).
    template <typename A>
    __device__ __forceinline__ uint32_t spawn_sum(uint32_t f, uint32_t q, const A& a) {
        child_sum += 1;
        return spawn_impl(f, q, tid, kResSum, a);
    }

    template <typename A>
    __device__ __forceinline__ uint32_t spawn_impl(uint32_t f, uint32_t q, uint32_t par, uint32_t roff, const A& a) {
        const uint32_t id = spawn_record<P>(*dev, w, me, tid, f, q, par, roff, a);
        if (par != kNoParent) child_count += 1;
        return id;
    }

    // taskwait: returns true when the task must return now (it has children to wait for);
    // the task resumes at `next_state`, re-enqueued on queue `q_after`. The record (header,
    // and the frame words of a task held in LDS) is written after the round, together with
    // the other suspended tasks of the wave.
    // (reference: __gtap_set_state_for_join, gtap_runtime_thread.cuh:672-696)
    __device__ __forceinline__ bool taskwait(uint32_t next_state, uint32_t q_after) {
        static_assert(!NoTaskwait<P>::value, "a program declared kNoTaskwait calls taskwait");
        if (q_after >= (uint32_t)Q) {
            GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_INVALID_QUEUE_IDX_AFTER_JOIN, tid, q_after, q_after, Q);
            q_after = 0;
        }
        if (child_sum != 0) {
            // children of an additive join add into the join word; only taskwait_sum waits for them
            GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_INVALID_TASKWAIT, tid, -1, child_sum, 0);
            child_sum = 0;
        }
        if (child_count == 0) {
            state = next_state;
            return false;
        }
        state = next_state;
        queue = q_after;
        join_count = child_count;
        child_count = 0;
        outcome = OUT_SUSPENDED;
        return true;
    }

    // taskwait over children spawned with spawn_sum: when the task resumes at `next_state`,
    // join_sum() is the sum of their results (0 when there were none). Every child joined
    // here must be a spawn_sum child.
    __device__ __forceinline__ bool taskwait_sum(uint32_t next_state, uint32_t q_after) {
        static_assert(!NoTaskwait<P>::value, "a program declared kNoTaskwait calls taskwait");
        if (q_after >= (uint32_t)Q) {
            GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_INVALID_QUEUE_IDX_AFTER_JOIN, tid, q_after, q_after, Q);
            q_after = 0;
        }
        state = next_state;
        jsum = 0;
        if (child_count == 0) return false;
        if (child_sum != child_count || child_count > kSumMaxCount || func > kSumMaxFunc || next_state > kSumMaxState) {
            GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_INVALID_TASKWAIT, tid, -1, child_sum, child_count);
        }
        queue = q_after;
        join_count = child_count & kSumMaxCount;
        child_count = 0;
        child_sum = 0;
        sum_join = true;
        outcome = OUT_SUSPENDED;
        return true;
    }
    __device__ __forceinline__ int64_t join_sum() const { return jsum; }

    // return value: written into the parent's result slot, or (a child of an additive join)
    // carried by the join atomic
    template <typename T>
    __device__ __forceinline__ void ret(T v) {
        if (parent != kNoParent) {
            if (res_off == kResSum) {
                if constexpr (std::is_integral<T>::value) {
                    const int64_t x = (int64_t)v;
                    if (x >= kSumMaxValue || x <= -kSumMaxValue)
                        GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_INVALID_ARGUMENT, tid, -1, (int)(x >> 32), (int)x);
                    join_delta = ((uint64_t)x << 16) - 1u;
                } else {
                    GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_INVALID_ARGUMENT, tid, -1, 0, 0);
                }
            } else {
                stT<T>(rec_ptr(*dev, parent) + res_off, v);
            }
        }
        finish();
    }
    __device__ __forceinline__ void finish() {
        if (child_count != 0) {
            // joined children spawned after the last taskwait would notify a finished
            // parent; such children must be spawned detached
            GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_INVALID_TASKWAIT, tid, -1, child_count, 0);
        }
        outcome = OUT_FINISHED;
    }
};

// ------------------------------------------------------------------ wave-cooperative tasks
// A program may name one EPAQ queue (P::kWaveQueue >= 0) whose tasks run on a whole
// wavefront: a batch taken from that queue holds one task and all 64 lanes enter
// P::run_wave with the same task. Spawn, taskwait and return are called by all lanes with
// uniform arguments; lane 0 does the bookkeeping.
template <class P>
struct WaveTask {
    static constexpr int Q = P::kQueues;
    const GtapDev* dev;
    WaveCtxOf<P>* w;
    const typename P::Args* args;
    typename P::WaveShared* sh;
    uint8_t* rec;
    uint32_t me, tid, func, state, queue, parent, res_off, child_count, outcome;
    int lane;
    uint32_t frame[P::kFrameWords];

    __device__ __forceinline__ uint32_t get_state() const { return state; }
    template <typename T>
    __device__ __forceinline__ T arg(int byte_off) const {
        if constexpr (sizeof(T) == 4) {
            return __builtin_bit_cast(T, frame[byte_off / 4]);
        } else {
            const uint64_t v = (uint64_t)frame[byte_off / 4] | ((uint64_t)frame[byte_off / 4 + 1] << 32);
            return __builtin_bit_cast(T, v);
        }
    }
    template <typename T>
    __device__ __forceinline__ T frame_ld(int byte_off) const { return ldT<T>(rec + kHdrBytes + byte_off); }
    template <typename T>
    __device__ __forceinline__ void frame_st(int byte_off, T v) {
        if (lane == 0) stT<T>(rec + kHdrBytes + byte_off, v);
    }
    template <typename A>
    __device__ __forceinline__ void spawn(uint32_t f, uint32_t q, int res_off_in_frame, const A& a) {
        if (lane == 0) spawn_record<P>(*dev, w, me, tid, f, q, tid, kHdrBytes + (uint32_t)res_off_in_frame, a);
        child_count += 1;
    }
    template <typename A>
    __device__ __forceinline__ void spawn_detached(uint32_t f, uint32_t q, const A& a) {
        if (lane == 0) spawn_record<P>(*dev, w, me, tid, f, q, kNoParent, 0u, a);
    }
    // Every calling lane spawns a detached child of its own (function, queue and arguments
    // per lane); any subset of the lanes may call, in divergent control flow too.
    template <typename A>
    __device__ __forceinline__ void spawn_detached_lane(uint32_t f, uint32_t q, const A& a) {
        spawn_record<P>(*dev, w, me, tid, f, q, kNoParent, 0u, a);
    }
    __device__ __forceinline__ bool taskwait(uint32_t next_state, uint32_t q_after) {
        static_assert(!NoTaskwait<P>::value, "a program declared kNoTaskwait calls taskwait");
        if (q_after >= (uint32_t)Q) q_after = 0;
        if (child_count == 0) {
            state = next_state;
            return false;
        }
        if (lane == 0)
            st128(pool_rsrc(dev->recs), tid * dev->stride,
                  u32x4{parent, res_off, child_count, hdr_word0(func, next_state, q_after)});
        child_count = 0;
        outcome = OUT_SUSPENDED;
        return true;
    }
    template <typename T>
    __device__ __forceinline__ void ret(T v) {
        if (lane == 0 && parent != kNoParent) {
            if (res_off == kResSum) GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_INVALID_TASKWAIT, tid, -1, 0, 0);  // lane tasks only
            else stT<T>(rec_ptr(*dev, parent) + res_off, v);
        }
        finish();
    }
    __device__ __forceinline__ void finish() {
        if (child_count != 0 && lane == 0) GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_INVALID_TASKWAIT, tid, -1, child_count, 0);
        outcome = OUT_FINISHED;
    }
};

// Runs the wave task `tid` on all lanes; returns its outcome (uniform).
template <class P>
__device__ __forceinline__ uint32_t run_wave_task(const GtapDev& dev, const typename P::Args& args,
                                                  WaveCtxOf<P>* w, void* wave_shared, uint32_t me,
                                                  uint32_t tid, uint32_t& parent_out) {
    if constexpr (HasWaveQueue<P>::value >= 0) {
        WaveTask<P> W;
        W.dev = &dev;
        W.w = w;
        W.args = &args;
        W.sh = reinterpret_cast<typename P::WaveShared*>(wave_shared);
        W.me = me;
        W.lane = lane_id();
        W.tid = tid;
        W.rec = rec_ptr(dev, tid);
        const u32x4 h = ld128(pool_rsrc(dev.recs), tid * dev.stride);
        W.func = hdr_func(h.w);
        W.state = hdr_state(h.w);
        W.queue = hdr_queue(h.w);
        W.parent = h.x;
        W.res_off = h.y;
        W.child_count = 0;
        W.outcome = OUT_RUNNING;
        const uint64_t* fp = reinterpret_cast<const uint64_t*>(W.rec + kHdrBytes);
#pragma unroll
        for (int i = 0; i < P::kFrameWords / 2; ++i) {
            const uint64_t v = ld64(fp + i);
            W.frame[2 * i] = (uint32_t)v;
            W.frame[2 * i + 1] = (uint32_t)(v >> 32);
        }
        if (W.func == 0 || W.func > P::kMaxFunc) {
            if (W.lane == 0) GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_FUNCTION, tid, W.queue, W.func, P::kMaxFunc);
            W.outcome = OUT_FINISHED;
        } else {
            P::run_wave(W);
        }
        if (W.outcome == OUT_RUNNING) {
            if (W.lane == 0) GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_TASKWAIT, tid, W.queue, W.state, 0);
            W.outcome = OUT_FINISHED;
        }
        parent_out = W.parent;
        return W.outcome;
    } else {
        (void)dev; (void)args; (void)w; (void)wave_shared; (void)me; (void)tid;
        parent_out = kNoParent;
        return OUT_FINISHED;
    }
}

// ------------------------------------------------------------------ scheduler

struct LoopStats {
    uint64_t tasks, rounds, pops, steals, attempts, jobs, published, suspended, ring_taken;
};

template <class P>
__device__ __forceinline__ void load_task(Lane<P>& L, uint32_t id) {
    const GtapDev& dev = *L.dev;
    L.tid = id;
    L.rec = rec_ptr(dev, id);
    const uint32_t off = id * dev.stride;
    const u32x4 h = ld128(L.rs, off);
    if (h.y & kSumLayout) {
        // resumed after an additive join: the join word holds the children's sum
        L.func = (h.y >> 24) & kSumMaxFunc;
        L.state = (h.y >> 16) & kSumMaxState;
        L.queue = (h.z >> 12) & 0xFu;
        L.res_off = h.y & 0xFFFFu;
        L.jsum = (int64_t)(((uint64_t)h.w << 32) | (uint64_t)h.z) >> 16;
    } else {
        L.func = hdr_func(h.w);
        L.state = hdr_state(h.w);
        L.queue = hdr_queue(h.w);
        L.res_off = h.y;
        L.jsum = 0;
    }
    L.parent = h.x;
    L.child_count = 0;
    L.child_sum = 0;
    L.sum_join = false;
    L.join_delta = ~0ull;
    L.outcome = OUT_RUNNING;
    L.lds = false;
    frame_words_ld<CtxShape<P>::kFWr>(L.rs, off, L.frame);
}

// A task held in LDS: its descriptor { word0, parent, result offset, frame words }.
template <class P>
__device__ __forceinline__ void load_task_lds(Lane<P>& L, uint32_t id, const uint32_t* d) {
    L.tid = id;
    L.rec = rec_ptr(*L.dev, id);
    const uint32_t w0 = vld(&d[0]);
    L.func = hdr_func(w0);
    L.state = hdr_state(w0);
    L.queue = hdr_queue(w0);
    L.parent = vld(&d[1]);
    L.res_off = vld(&d[2]);
    L.child_count = 0;
    L.child_sum = 0;
    L.sum_join = false;
    L.jsum = 0;
    L.join_delta = ~0ull;
    L.outcome = OUT_RUNNING;
    L.lds = true;
#pragma unroll
    for (int i = 0; i < P::kFrameWords; ++i) L.frame[i] = vld(&d[3 + i]);
#pragma unroll
    for (int i = P::kFrameWords; i < CtxShape<P>::kFWr; ++i) L.frame[i] = 0u;
}

// Job j of the table becomes a root task: its anchor record (join = 1, the root) and the
// root record with the job's arguments; the job is marked running. Returns the root id.
// Called by the lanes that start a job each.
template <class P>
__device__ __forceinline__ uint32_t start_job(const GtapDev& dev, WaveCtxOf<P>* w, uint32_t me, uint32_t j) {
    const uint32_t slot = j % dev.job_ring;
    gtap_job* job = dev.jobs + slot;
    const uint32_t func = ld32_sys(&job->func);
    const uint32_t q = ld32_sys(&job->queue);
    const uint32_t anchor = dev.nrecs + slot;
    const uint32_t id = alloc_id(dev, w, me);
    uint8_t* arec = rec_ptr(dev, anchor);
    st64(reinterpret_cast<uint64_t*>(arec), hdr_lo(kNoParent, 0u));
    st64(reinterpret_cast<uint64_t*>(arec + kHdrJoin), hdr_hi(1u, hdr_word0(kAnchorFunc, 0, 0)));  // join = 1 (the root)
    st64(reinterpret_cast<uint64_t*>(arec + kHdrBytes), 0ull);
    uint8_t* rrec = rec_ptr(dev, id);
    st64(reinterpret_cast<uint64_t*>(rrec), hdr_lo(anchor, kHdrBytes));
    st64(reinterpret_cast<uint64_t*>(rrec + kHdrJoin), hdr_hi(0u, hdr_word0(func, 0, q < (uint32_t)P::kQueues ? q : 0)));
    const uint64_t* ja = reinterpret_cast<const uint64_t*>(job->args);
    uint64_t* ra = reinterpret_cast<uint64_t*>(rrec + kHdrBytes);
    constexpr int kArgWords = (P::kFrameWords / 2) < 8 ? (P::kFrameWords / 2) : 8;
#pragma unroll
    for (int i = 0; i < kArgWords; ++i) st64(ra + i, ld64_sys(ja + i));
    st64_sys(&job->t_start, now_ticks());
    st32_sys(&job->status, (uint32_t)GTAP_JOB_RUNNING);
    return id;
}

// Admit up to 64 jobs from the job table as root tasks. Returns the number admitted; lane
// i < count receives the root id in `tid_out`.
template <class P>
__device__ __forceinline__ uint32_t admit_jobs(const GtapDev& dev, WaveCtxOf<P>* w, uint32_t me,
                                              uint32_t& tid_out, bool& idle_registered, uint32_t& seen_next) {
    const int lane = lane_id();
    uint32_t avail = dev.njobs;
    if (dev.term_mode == GTAP_TERM_SERVE) avail = uniform((uint32_t)ld32(dev.ctl + CTL_HOST_AVAIL * 32));
    // the admission counter only grows: once it was seen at `avail`, nothing is left to admit
    // and the shared counter's line is not read again
    if (seen_next >= avail) return 0;
    int32_t* nextp = dev.ctl + CTL_JOB_NEXT * 32;
    uint64_t* wordp = reinterpret_cast<uint64_t*>(nextp);  // {admission counter, reserved units}
    const uint64_t word = uniform(ld64(wordp));
    const uint32_t next = (uint32_t)word;
    seen_next = next;
    if (next >= avail) return 0;
    uint32_t k = avail - next;
    if (k > kWave) k = kWave;
    // admission: cap on live jobs, and (served) the reservation ledger
    const uint32_t live = uniform((uint32_t)ld32(dev.ctl + CTL_LIVE_JOBS * 32));
    if (live >= dev.max_live_jobs) return 0;
    if (k > dev.max_live_jobs - live) k = dev.max_live_jobs - live;
    uint32_t units = 0;
    if (dev.admit_budget > 0) {
        // the longest run of the next jobs whose reservations fit beside those of the running jobs
        const uint32_t ledger = (uint32_t)(word >> 32);
        const uint32_t room = dev.admit_budget > ledger ? dev.admit_budget - ledger : 0u;
        uint32_t u = (uint32_t)lane < k ? ld32(dev.job_units + (next + (uint32_t)lane) % dev.job_ring) : 0u;
        uint32_t pre = u;  // inclusive prefix sum over the lanes
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t t = (uint32_t)__shfl_up((int)pre, (unsigned)d, kWave);
            if (lane >= d) pre += t;
        }
        const uint32_t m = (uint32_t)popc(ballot((uint32_t)lane < k && pre <= room));
        if (m == 0) {
            // held back: the job at the head waits for running jobs to release their ids
            if (lane == 0 && me == next % dev.nworkers) atomic_add_noret(dev.ctl + CTL_ADMIT_HOLDS * 32, 1);
            return 0;
        }
        k = m;
        units = shfl(pre, (int)m - 1);
    }
    if (dev.term_mode == GTAP_TERM_QUIESCE && idle_registered) {
        if (lane == 0) atomic_add(dev.ctl + CTL_IDLE * 32, -1);
        idle_registered = false;
    }
    int ok = 0;
    if (lane == 0) {
        // the live count is raised before the admission counter moves (and lowered again if
        // the claim fails): a served run's stop rule reads the counter, then the live count,
        // and must not see the jobs admitted but not yet live
        atomic_add(dev.ctl + CTL_LIVE_JOBS * 32, (int32_t)k);
        complete_stores();
        uint64_t expected = word;
        ok = atomic_cas(wordp, expected, word + (uint64_t)k + ((uint64_t)units << 32)) ? 1 : 0;
        if (!ok) atomic_add_noret(dev.ctl + CTL_LIVE_JOBS * 32, -(int32_t)k);
    }
    ok = shfl((int32_t)ok, 0);
    if (!ok) return 0;
    if ((uint32_t)lane < k) tid_out = start_job<P>(dev, w, me, next + (uint32_t)lane);
    complete_stores();
    return k;
}

// Jobs [0, dev.static_jobs) are handed out by index at launch start: worker j starts job j
// without touching the admission counter, which the reset sets to static_jobs. Lane 0 takes
// the root; returns the batch size (1 or 0).
template <class P>
__device__ __forceinline__ uint32_t admit_static(const GtapDev& dev, WaveCtxOf<P>* w, uint32_t me, uint32_t& tid_out) {
    if (me >= dev.static_jobs) return 0;
    if (lane_id() == 0) tid_out = start_job<P>(dev, w, me, me);
    complete_stores();
    return 1;
}

// A job's root finished: publish its result to the job table.
template <class P>
__device__ __forceinline__ void complete_job(const GtapDev& dev, uint32_t anchor) {
    const uint32_t slot = anchor - dev.nrecs;
    gtap_job* job = dev.jobs + slot;
    if (dev.admit_budget > 0) {
        // release the job's reservation (read before the slot is handed back to the host)
        const uint64_t units = ld32(dev.job_units + slot);
        atomic_add_noret(reinterpret_cast<uint64_t*>(dev.ctl + CTL_JOB_NEXT * 32), (uint64_t)0 - (units << 32));
    }
    const uint64_t res = ld64(reinterpret_cast<const uint64_t*>(rec_ptr(dev, anchor) + kHdrBytes));
    st64_sys(reinterpret_cast<uint64_t*>(&job->result), res);
    st64_sys(&job->t_end, now_ticks());
    complete_stores();
    st32_sys(&job->status, (uint32_t)GTAP_JOB_DONE);
    atomic_add(dev.ctl + CTL_LIVE_JOBS * 32, -1);
    const int32_t done = atomic_add(dev.ctl + CTL_JOBS_DONE * 32, 1) + 1;
    if (dev.term_mode == GTAP_TERM_JOBS && (uint32_t)done == dev.njobs) st32(dev.ctl + CTL_DONE * 32, 1);
}

// A random victim other than `me` for one probe lane. With xcd_steal, even lanes pick one on
// this wave's XCD: workgroups are dealt to the 8 XCDs round-robin, so workgroup 8 k + x runs
// on XCD x; 4 waves per workgroup.
__device__ __forceinline__ uint32_t pick_victim(const GtapDev& dev, uint32_t me, uint32_t& rng, int lane) {
    uint32_t victim = xorshift32(rng) % (dev.nworkers - 1);
    if (victim >= me) victim += 1;
    if (dev.xcd_steal && (lane & 1) == 0 && dev.nworkers >= 64 && (dev.nworkers & 3u) == 0) {
        const uint32_t nb = dev.nworkers >> 2, x = (me >> 2) & 7u;
        const uint32_t nbx = (nb - x + 7u) >> 3;  // workgroups on XCD x
        const uint32_t vw = (((xorshift32(rng) % nbx) << 3) + x) * 4u + (victim & 3u);
        if (vw != me && vw < dev.nworkers) victim = vw;
    }
    return victim;
}

// With several EPAQ queues, one probe of every queue at once: lane l probes queue
// (cq + l) % Q of a random victim (probe_lanes victims per queue, at most 64 probes), and one
// CAS claims up to half (<= steal_batch; 1 on the wave queue) of a non-empty one. The thief
// finds work in one round trip instead of one per queue. Returns the number stolen into
// `tid_out` and the queue in `q_out`.
template <class P>
__device__ __forceinline__ uint32_t try_steal_any(const GtapDev& dev, uint32_t me, uint32_t cq, uint32_t& rng,
                                                  uint32_t& tid_out, bool& idle_registered, LoopStats& st,
                                                  uint32_t& q_out) {
    constexpr uint32_t Q = (uint32_t)P::kQueues;
    const int lane = lane_id();
    if (dev.nworkers < 2) return 0;
    uint32_t probes = dev.probe_lanes * Q;
    if (probes > (uint32_t)kWave) probes = kWave;
    uint32_t victim = 0, q = (cq + (uint32_t)lane) % Q;
    uint64_t wv = 0;
    int32_t sz = 0;
    if ((uint32_t)lane < probes) {
        victim = pick_victim(dev, me, rng, lane);
        wv = ld64(dq_word_ptr(dev, q, victim));
        sz = (int32_t)((uint32_t)wv - (uint32_t)(wv >> 32));
    }
    uint64_t cand = ballot(sz > 0);
    if (cand == 0) return 0;
    const uint32_t rot = uniform(xorshift32(rng)) & 63u;
    const uint64_t rotated = rot ? ((cand >> rot) | (cand << (64 - rot))) : cand;
    const int src = (first_lane(rotated) + (int)rot) & 63;
    victim = shfl(victim, src);
    wv = shfl(wv, src);
    sz = shfl(sz, src);
    q = uniform(shfl(q, src));
    uint32_t k = (uint32_t)(sz + 1) / 2;
    const uint32_t batch_cap = ((int)q == HasWaveQueue<P>::value) ? 1u : dev.steal_batch;
    if (k > batch_cap) k = batch_cap;
    if (k == 0) k = 1;
    if (dev.term_mode == GTAP_TERM_QUIESCE && idle_registered) {
        if (lane == 0) atomic_add(dev.ctl + CTL_IDLE * 32, -1);
        idle_registered = false;
    }
    int ok = 0;
    if (lane == 0) {
        uint64_t expected = wv;
        ok = atomic_cas(dq_word_ptr(dev, q, victim), expected, wv + ((uint64_t)k << 32)) ? 1 : 0;
    }
    ok = shfl((int32_t)ok, 0);
    st.attempts += 1;
    if (!ok) return 0;
    const uint32_t t = (uint32_t)(wv >> 32);
    if ((uint32_t)lane < k) tid_out = ld32(dq_slot_ptr(dev, q, victim, t + (uint32_t)lane));
    st.steals += 1;
    q_out = q;
    return k;
}

// Try to steal from random victims on queue q. Returns the number of ids stolen into `tid_out`.
template <class P>
__device__ __forceinline__ uint32_t try_steal(const GtapDev& dev, uint32_t me, uint32_t q, uint32_t& rng,
                                              uint32_t& tid_out, bool& idle_registered, LoopStats& st,
                                              uint32_t max_items = kWave) {
    uint32_t batch_cap = ((int)q == HasWaveQueue<P>::value) ? 1u : dev.steal_batch;
    if (batch_cap > max_items) batch_cap = max_items;
    const int lane = lane_id();
    if (dev.nworkers < 2) return 0;
    uint32_t victim = 0;
    uint64_t wv = 0;
    int32_t sz = 0;
    if ((uint32_t)lane < dev.probe_lanes) {
        victim = pick_victim(dev, me, rng, lane);
        wv = ld64(dq_word_ptr(dev, q, victim));
        sz = (int32_t)((uint32_t)wv - (uint32_t)(wv >> 32));
    }
    uint64_t cand = ballot(sz > 0);
    if (cand == 0) return 0;
    // rotate the candidate mask by a random amount so thieves spread over victims
    const uint32_t rot = uniform(xorshift32(rng)) & 63u;
    const uint64_t rotated = rot ? ((cand >> rot) | (cand << (64 - rot))) : cand;
    const int src = (first_lane(rotated) + (int)rot) & 63;
    victim = shfl(victim, src);
    wv = shfl(wv, src);
    sz = shfl(sz, src);
    uint32_t k = (uint32_t)(sz + 1) / 2;
    if (k > batch_cap) k = batch_cap;
    if (k == 0) k = 1;
    if (dev.term_mode == GTAP_TERM_QUIESCE && idle_registered) {
        if (lane == 0) atomic_add(dev.ctl + CTL_IDLE * 32, -1);
        idle_registered = false;
    }
    int ok = 0;
    if (lane == 0) {
        uint64_t expected = wv;
        ok = atomic_cas(dq_word_ptr(dev, q, victim), expected, wv + ((uint64_t)k << 32)) ? 1 : 0;
    }
    ok = shfl((int32_t)ok, 0);
    st.attempts += 1;
    if (!ok) return 0;
    const uint32_t t = (uint32_t)(wv >> 32);
    if ((uint32_t)lane < k) tid_out = ld32(dq_slot_ptr(dev, q, victim, t + (uint32_t)lane));
    st.steals += 1;
    return k;
}

// Pop up to 64 (max_items) of this wave's newest published tasks from queue q. The queue must
// hold no private entries (its tail equals its bottom): both move to the new bottom.
template <class P>
__device__ __forceinline__ uint32_t try_pop(const GtapDev& dev, WaveCtxOf<P>* w, uint32_t me, uint32_t q,
                                            uint32_t& tid_out, uint32_t max_items = kWave) {
    const int lane = lane_id();
    uint64_t* wp = dq_word_ptr(dev, q, me);
    // A failed CAS means a thief took tasks, so this ends; the owner leaves with a task or an
    // empty deque. Only the owner pushes, so the deque stays empty while the owner is idle
    // (idle waves do not pop again, and quiescence may count them idle).
    while (true) {
        uint64_t wv = 0;
        if (lane == 0) wv = ld64(wp);
        wv = shfl(wv, 0);
        const uint32_t t = (uint32_t)(wv >> 32), b = (uint32_t)wv;
        const int32_t sz = (int32_t)(b - t);
        if (sz <= 0) return 0;
        uint32_t cap = ((int)q == HasWaveQueue<P>::value) ? 1u : (uint32_t)kWave;
        if (cap > max_items) cap = max_items;
        const uint32_t k = (uint32_t)sz < cap ? (uint32_t)sz : cap;
        int ok = 0;
        if (lane == 0) {
            uint64_t expected = wv;
            ok = atomic_cas(wp, expected, ((uint64_t)t << 32) | (uint64_t)(b - k)) ? 1 : 0;
        }
        ok = shfl((int32_t)ok, 0);
        if (ok) {
            if ((uint32_t)lane < k) tid_out = ld32(dq_slot_ptr(dev, q, me, b - k + (uint32_t)lane));
            vst(&w->bottom[q], b - k);
            vst(&w->tail[q], b - k);
            vst(&w->top_cache[q], t);
            return k;
        }
    }
}

// ------------------------------------------------------------------ other schedulers
// The alternatives the reference compares its batched work stealing with
// (runtime/detail/experimental/thread/gtap_runtime_thread_gq.cuh:484-621,
// gtap_runtime_thread_chaselev.cuh:516-712), kept for the worker-scaling comparison:
//   kSchedGQ  one global FIFO of task ids per EPAQ queue. A wave claims up to 64 ids from
//             its head with one CAS. A publishing wave reserves a range with a fetch-add,
//             writes the ids, then waits until every earlier range is committed and
//             advances the committed tail (in-order commit).
//   kSchedCL  the owner pops and thieves steal one task per CAS; a wave fills its batch
//             with up to 64 such single-item operations in a row.
enum : int { kSchedWS = 0, kSchedGQ = 1, kSchedCL = 2 };
// Work stealing with several EPAQ queues probes all of them in one steal attempt (try_steal_any).
constexpr bool kStealAnyQueue = true;
constexpr uint32_t kGqMargin = 1024;  // ring slots kept free beyond the head

// Control words of global queue q: head, reservation counter and committed tail, each on its
// own 128-byte line.
__device__ __forceinline__ uint64_t* gq_word(const GtapDev& dev, uint32_t q, uint32_t which) {
    return dev.gq_ctl + ((size_t)q * 3 + which) * 16;
}
__device__ __forceinline__ uint32_t* gq_slot(const GtapDev& dev, uint32_t q, uint64_t pos) {
    return dev.gq_slots + (size_t)q * dev.gq_cap + (pos & (uint64_t)(dev.gq_cap - 1));
}

// Claim up to 64 (1 for the wave queue) committed ids from the head of global queue q.
template <class P>
__device__ __forceinline__ uint32_t gq_pop(const GtapDev& dev, uint32_t q, uint32_t& tid_out, bool& idle_registered,
                                           LoopStats& st) {
    const int lane = lane_id();
    const uint32_t cap = ((int)q == HasWaveQueue<P>::value) ? 1u : (uint32_t)kWave;
    uint64_t* headp = gq_word(dev, q, 0);
    uint64_t* tailp = gq_word(dev, q, 2);
    for (int tries = 0; tries < 8; ++tries) {
        uint64_t h = 0, tl = 0;
        if (lane == 0) {
            h = ld64(headp);
            tl = ld64(tailp);
        }
        h = shfl(h, 0);
        tl = shfl(tl, 0);
        if ((int64_t)(tl - h) <= 0) return 0;
        const uint32_t k = (tl - h) < (uint64_t)cap ? (uint32_t)(tl - h) : cap;
        if (dev.term_mode == GTAP_TERM_QUIESCE && idle_registered) {
            if (lane == 0) atomic_add(dev.ctl + CTL_IDLE * 32, -1);
            idle_registered = false;
        }
        int ok = 0;
        if (lane == 0) {
            uint64_t expected = h;
            ok = atomic_cas(headp, expected, h + k) ? 1 : 0;
        }
        ok = shfl((int32_t)ok, 0);
        st.attempts += 1;
        if (ok) {
            if ((uint32_t)lane < k) tid_out = ld32(gq_slot(dev, q, h + (uint32_t)lane));
            st.steals += 1;
            return k;
        }
    }
    return 0;
}

// Append the ids at positions [b, t) of this wave's own slot array for queue q (all written
// and completed) to global queue q: reserve, copy, wait for the earlier reservations to
// commit, commit. The own array then holds nothing published: its top moves to t.
template <class P>
__device__ __forceinline__ void gq_push(const GtapDev& dev, WaveCtxOf<P>* w, uint32_t me, uint32_t q, uint32_t b,
                                        uint32_t t) {
    const int lane = lane_id();
    const uint32_t n = t - b;
    uint64_t base = 0;
    int overflow = 0;
    if (lane == 0) {
        base = atomic_add(gq_word(dev, q, 1), (uint64_t)n);
        const uint64_t head = ld64(gq_word(dev, q, 0));
        overflow = (base + n - head > (uint64_t)(dev.gq_cap - kGqMargin)) ? 1 : 0;
        if (overflow) GTAP_DEVICE_ERROR(dev, GTAP_ERROR_QUEUE_OVERFLOW, -1, q, (int)(base + n - head), dev.gq_cap - kGqMargin);
    }
    base = shfl(base, 0);
    overflow = shfl((int32_t)overflow, 0);
    if (!overflow) {
        for (uint32_t j = (uint32_t)lane; j < n; j += kWave) st32(gq_slot(dev, q, base + j), ld32(dq_slot_ptr(dev, q, me, b + j)));
        complete_stores();  // the ids are in the ring before the range is committed
        if (lane == 0) {
            // in-order commit: every earlier reservation commits in bounded time (its owner
            // only copies ids between reserving and committing); an abort ends the wait
            uint64_t* tailp = gq_word(dev, q, 2);
            for (uint32_t spin = 0;; ++spin) {
                if (ld64(tailp) == base) {
                    atomic_add_noret(tailp, (uint64_t)n);
                    break;
                }
                if ((spin & 255u) == 255u && ld32(dev.ctl + CTL_ABORT * 32)) break;
                if (spin > (1u << 24)) {
                    GTAP_DEVICE_ERROR(dev, GTAP_ERROR_WATCHDOG, -1, q, (int)base, (int)n);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
    if (lane == 0) {
        // nothing of [b, t) stays in the own array: its top and bottom move to t
        st64(dq_word_ptr(dev, q, me), ((uint64_t)t << 32) | (uint64_t)t);
        vst(&w->top_cache[q], t);
    }
}

// Single-item Chase-Lev pops: lane k receives the k-th task popped, one CAS per task.
template <class P>
__device__ __forceinline__ uint32_t cl_pop_batch(const GtapDev& dev, WaveCtxOf<P>* w, uint32_t me, uint32_t q,
                                                 uint32_t& tid_out) {
    const int lane = lane_id();
    const uint32_t cap = ((int)q == HasWaveQueue<P>::value) ? 1u : (uint32_t)kWave;
    uint32_t k = 0;
    while (k < cap) {
        uint32_t one = 0;
        if (!try_pop<P>(dev, w, me, q, one, 1u)) break;
        one = shfl(one, 0);
        if ((uint32_t)lane == k) tid_out = one;
        ++k;
    }
    return k;
}

__device__ __forceinline__ void backoff_sleep(uint32_t level) {
    switch (level) {
        case 0: break;
        case 1: __builtin_amdgcn_s_sleep(1); break;
        case 2: __builtin_amdgcn_s_sleep(2); break;
        case 3: __builtin_amdgcn_s_sleep(4); break;
        case 4: __builtin_amdgcn_s_sleep(8); break;
        case 5: __builtin_amdgcn_s_sleep(16); break;
        default: __builtin_amdgcn_s_sleep(32); break;
    }
}

// Prof: the kernel records the timeline (gtap_set_profiling). It is a separate build of the
// kernel, because the timeline's loop state costs registers every launch would pay for.
template <class P, int Sched, bool Prof>
__device__ __forceinline__ void thread_scheduler(const GtapDev& dev, const typename P::Args& args, WaveCtxOf<P>* w,
                                 void* wave_shared) {
    constexpr int Q = P::kQueues;
    constexpr int WQ = HasWaveQueue<P>::value;
    const int lane = lane_id();
    const uint32_t me = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t t_begin = now_ticks();

    if (lane == 0) {
        vst(&w->free_top, 0u);
        vst(&w->priv_top, 0u);
        vst(&w->cur_queue, 0u);
        vst(&w->spawns, 0u);
        vst(&w->refills, 0u);
        vst(&w->overflowed, 0u);
        vst(&w->ring_spilled, 0u);
        vst(&w->nstage, 0u);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            vst(&w->tail[q], 0u);
            vst(&w->bottom[q], 0u);
            vst(&w->top_cache[q], 0u);
            vst(&w->valid[q], (uint64_t)0);
            vst(&w->inlds[q], (uint64_t)0);
        }
        // a worker that starts a job by index was counted as started by the reset
        if (dev.term_mode == GTAP_TERM_QUIESCE && me >= dev.static_jobs) atomic_add(dev.ctl + CTL_STARTED * 32, 1);
    }
    __builtin_amdgcn_wave_barrier();
    // the worker's first ids are a fixed range (the shared fresh-id counter starts above
    // every worker's range), so starting needs no shared counter
    if (dev.static_ids) {
        for (uint32_t i = (uint32_t)lane; i < dev.static_ids; i += kWave) vst(&w->free_ids[i], 64u + me * dev.static_ids + i);
        if (lane == 0) vst(&w->free_top, dev.static_ids);
        __builtin_amdgcn_wave_barrier();
    }

    Lane<P> L;
    L.dev = &dev;
    L.w = w;
    L.args = &args;
    L.me = me;
    L.rs = pool_rsrc(dev.recs);
    constexpr int QL = CtxShape<P>::kQL;

    uint32_t rng = 0x9E3779B9u ^ (me * 0x85EBCA6Bu) ^ ((uint32_t)lane * 0xC2B2AE35u) ^ (uint32_t)t_begin;
    if (rng == 0) rng = 1;
    uint32_t n = 0;        // tasks in the current batch (wave-uniform)
    uint32_t my_tid = 0;   // lane's task when lane < n
    bool my_lds = false;   // the lane's task is held in LDS: desc[batch_q][my_slot]
    uint32_t my_slot = 0;
    int batch_q = -1;      // queue the batch came from (-1: admitted job roots)
    bool idle_registered = false;
    uint32_t idle_iter = 0;
    uint32_t busy_iter = 0;
    // last value of the shared admission counter this wave read (the jobs handed out by
    // index are taken)
    uint32_t seen_next = dev.jobs != nullptr ? dev.static_jobs : 0u;
    LoopStats st = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // timeline profiling (lane 0, in the Prof build when dev.prof_max > 0)
    uint32_t ph_idx = 0, pw_idx = 0;
    bool had_tasks = false;
    // round-phase cycle sums (Prof build): wave-uniform, so they live in scalar registers
    uint64_t phase_acc[PH_ROUNDS] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t phase_t = Prof ? __builtin_amdgcn_s_memtime() : 0;
    auto phase_mark = [&](int k) {
        if constexpr (Prof) {
            const uint64_t now = __builtin_amdgcn_s_memtime();
            phase_acc[k] += now - phase_t;
            phase_t = now;
        }
    };

    uint64_t host_poll_last = 0;
    const bool host_poller = dev.term_mode == GTAP_TERM_SERVE && me == 0;
    // this wave's group epoch of the overflow rings when it last found them all empty; the
    // rings start a launch empty (the reset drains them), so the epoch read at the start will do
    uint32_t ring_seen = 0;
    if (Sched != kSchedGQ && dev.ovf_cells != nullptr) {
        if (lane == 0) ring_seen = ld32(dev.ovf_epoch + (me & 63u) * 32u);
        ring_seen = uniform(shfl(ring_seen, 0));
    }
    if (dev.jobs != nullptr) {
        n = admit_static<P>(dev, w, me, my_tid);
        st.jobs += n;
    }

    while (true) {
        if (host_poller && lane == 0) host_poll(dev, host_poll_last);
        if (n == 0) {
            // ---- acquire work: own deque, then the job table, then steal (records in HBM)
            my_lds = false;
            const uint32_t cq = uniform(vld(&w->cur_queue));
#pragma unroll
            for (int i = 0; i < Q && n == 0 && idle_iter == 0 && Sched != kSchedGQ; ++i) {
                const uint32_t q = (cq + (uint32_t)i) % (uint32_t)Q;
                if constexpr (Sched == kSchedCL) n = cl_pop_batch<P>(dev, w, me, q, my_tid);
                else n = try_pop<P>(dev, w, me, q, my_tid);
                if (n) {
                    vst(&w->cur_queue, q);
                    st.pops += 1;
                    batch_q = (int)q;
                }
            }
            if (n == 0 && dev.jobs != nullptr) {
                n = admit_jobs<P>(dev, w, me, my_tid, idle_registered, seen_next);
                st.jobs += n;
                batch_q = -1;
            }
            if (n == 0) {
                if constexpr (kStealAnyQueue && Sched == kSchedWS && Q > 1) {
                    uint32_t sq = 0;
                    n = try_steal_any<P>(dev, me, cq, rng, my_tid, idle_registered, st, sq);
                    if (n) batch_q = (int)sq;
                } else {
#pragma unroll
                    for (int i = 0; i < Q && n == 0; ++i) {
                        const uint32_t q = (cq + (uint32_t)i) % (uint32_t)Q;
                        if constexpr (Sched == kSchedGQ) n = gq_pop<P>(dev, q, my_tid, idle_registered, st);
                        else n = try_steal<P>(dev, me, q, rng, my_tid, idle_registered, st, Sched == kSchedCL ? 1u : (uint32_t)kWave);
                        if (n) batch_q = (int)q;
                    }
                }
            }
            // the overflow rings (tasks that full deques handed off), every fourth idle try
            // the overflow rings, when this wave's group epoch moved since it last found them empty
            bool rings_seen_empty = true;
            if (Sched != kSchedGQ && n == 0 && dev.ovf_cells != nullptr) {
                uint32_t ep = 0;
                if (lane == 0) ep = ld32(dev.ovf_epoch + (me & 63u) * 32u);
                ep = uniform(shfl(ep, 0));
                if (ep != ring_seen) {
                    bool all_empty = true;
#pragma unroll 1
                    for (int i = 0; i < Q && n == 0; ++i) {
                        const uint32_t q = (cq + (uint32_t)i) % (uint32_t)Q;
                        const uint32_t cap = ((int)q == WQ) ? 1u : (uint32_t)kWave;
                        const uint64_t got = ovf_pop(ovf_view(dev), q, cap, idle_registered ? 1 : 0);
                        const uint32_t code = uniform((uint32_t)got);
                        if (code & kOvfLeftIdle) idle_registered = false;
                        if (!(code & kOvfEmpty)) all_empty = false;
                        n = code & ~(kOvfLeftIdle | kOvfEmpty);
                        if (n) {
                            if ((uint32_t)lane < n) my_tid = (uint32_t)(got >> 32);
                            batch_q = (int)q;
                            st.ring_taken += n;
                        }
                    }
                    if (all_empty) ring_seen = ep;
                    else rings_seen_empty = false;
                }
            }
            if (n == 0) {
                if (Prof && had_tasks) {
                    had_tasks = false;
                    if (dev.prof_max && lane == 0) prof_stamp(dev.prof_having, nullptr, dev.prof_max, me, ph_idx, 0);
                }
                // ---- idle: termination, abort, watchdog, backoff. A wave counts itself idle
                // only while no overflow ring holds tickets: a wave that handed tasks to a
                // ring therefore sees them before it can count itself idle. It also waits
                // for kIdleRegisterAfter failed passes, so a short spell between steals costs
                // no atomics on the shared count. A wave not yet counted can only delay the
                // end of the launch (the count must reach every started wave), never cause it.
                if (dev.term_mode == GTAP_TERM_QUIESCE && !idle_registered && rings_seen_empty &&
                    idle_iter >= kIdleRegisterAfter) {
                    idle_registered = true;
                    int32_t c = 0;
                    int32_t started = 0;
                    if (lane == 0) {
                        c = atomic_add(dev.ctl + CTL_IDLE * 32, 1) + 1;
                        started = ld32(dev.ctl + CTL_STARTED * 32);
                    }
                    c = shfl(c, 0);
                    started = shfl(started, 0);
                    // every worker that has started is idle: no deque holds work, and a worker
                    // starting later begins empty
                    if (c == started) {
                        const uint32_t next = (uint32_t)ld32(dev.ctl + CTL_JOB_NEXT * 32);
                        if (next >= dev.njobs && lane == 0) st32(dev.ctl + CTL_DONE * 32, 1);
                    }
                }
                idle_iter += 1;
                int32_t flags = 0;
                if (lane == 0) {
                    flags = ld32(dev.ctl + CTL_DONE * 32);  // an abort sets it too
                    if (dev.term_mode == GTAP_TERM_SERVE && (idle_iter & 15u) == 0) {
                        const int32_t stop = ld32(dev.ctl + CTL_HOST_STOP * 32);
                        if (stop) {
                            const uint32_t avail = (uint32_t)ld32(dev.ctl + CTL_HOST_AVAIL * 32);
                            const uint32_t next = (uint32_t)ld32(dev.ctl + CTL_JOB_NEXT * 32);
                            complete_stores();  // the live count is read after the admission counter
                            const int32_t live = ld32(dev.ctl + CTL_LIVE_JOBS * 32);
                            if (next >= avail && live == 0) {
                                st32(dev.ctl + CTL_DONE * 32, 1);
                                flags |= 1;
                            }
                        }
                    }
                    if ((idle_iter & 63u) == 0 && now_ticks() - t_begin > dev.timeout_ticks) {
                        GTAP_DEVICE_ERROR(dev, GTAP_ERROR_WATCHDOG, -1, -1, (int)((now_ticks() - t_begin) / 100000u),
                                          (int)(dev.timeout_ticks / 100000u));
                        flags |= 2;
                    }
                }
                flags = shfl(flags, 0);
                if (flags) break;
                const uint32_t lvl = idle_iter < 8 ? 0u : (idle_iter < 64 ? 2u : (idle_iter < 256 ? 4u : 6u));
                backoff_sleep(lvl < dev.idle_level ? lvl : dev.idle_level);
                phase_mark(PH_IDLE);
                continue;
            }
            idle_iter = 0;
        }
        phase_mark(PH_ACQUIRE);

        // ---- execute one round: lane i runs task my_tid when i < n
        st.rounds += 1;
        st.tasks += n;
        const uint32_t round_n = n;
        if (Prof && dev.prof_max && lane == 0) {
            if (!had_tasks) prof_stamp(dev.prof_having, nullptr, dev.prof_max, me, ph_idx, 0);
            prof_stamp(dev.prof_working, dev.prof_counts, dev.prof_max, me, pw_idx, (int32_t)round_n);
        }
        if (Prof) had_tasks = true;
        uint32_t fin_tid = 0, fin_parent = kNoParent;
        uint64_t fin_delta = ~0ull;  // added to the parent's join word: -1, plus value << 16 (additive join)
        bool fin_sum = false;        // the parent waits in an additive join
        bool fin = false, sus = false;
        if (WQ >= 0 && batch_q == WQ) {
            // one wave-cooperative task on all lanes
            uint32_t wt = uniform(my_tid);
            if (wt == 0u || wt >= dev.nrecs) {
                if (lane == 0) GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_FUNCTION, wt, -1, wt, dev.nrecs);
                wt = 0u;
            }
            uint32_t par = kNoParent;
            const uint32_t outc = (wt != 0u) ? run_wave_task<P>(dev, args, w, wave_shared, me, wt, par) : OUT_FINISHED;
            fin = (lane == 0) && outc == OUT_FINISHED;
            fin_tid = wt;
            fin_parent = (wt != 0u && par < dev.nrecs + dev.nanchors) ? par : kNoParent;
        } else if ((uint32_t)lane < n) {
            if (my_tid == 0u || my_tid >= dev.nrecs) {
                // a slot never holds the sink or an anchor id; treat as a protocol error
                GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_FUNCTION, my_tid, -1, my_tid, dev.nrecs);
                my_tid = 0u;
            }
            if (my_lds && batch_q >= 0 && batch_q < QL) load_task_lds<P>(L, my_tid, &w->desc[batch_q < QL ? batch_q : 0][my_slot][0]);
            else load_task<P>(L, my_tid);
            if constexpr (Prof) {
                __asm__ volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");  // the loads' own time
                phase_mark(PH_LOAD);
            }
            if (my_tid == 0u) {
                L.outcome = OUT_FINISHED;
                L.parent = kNoParent;
            } else if (L.func == 0 || L.func > P::kMaxFunc) {
                GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_FUNCTION, L.tid, L.queue, L.func, P::kMaxFunc);
                L.outcome = OUT_FINISHED;
            } else {
                P::run(L);
            }
            if (L.outcome == OUT_RUNNING) {
                GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_TASKWAIT, L.tid, L.queue, L.state, 0);
                L.outcome = OUT_FINISHED;
            }
            fin = (L.outcome == OUT_FINISHED);
            sus = (L.outcome == OUT_SUSPENDED);
            fin_tid = L.tid;
            fin_parent = L.parent;
            fin_sum = L.res_off == kResSum;
            fin_delta = L.join_delta;
            if (fin_parent != kNoParent && fin_parent >= dev.nrecs + dev.nanchors) {
                GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_FUNCTION, fin_tid, -1, fin_parent, dev.nrecs + dev.nanchors);
                fin_parent = kNoParent;
            }
        }
        phase_mark(PH_RUN);

        // ---- epilogue: suspended records, joins, continuations, id release, next batch, publish
        {
            // records staged during the round (entries pushed out of the window), then the
            // tasks that suspended (up to 64 more), all written together
            const uint64_t sm = ballot(sus);
            const uint32_t base = uniform(vld(&w->nstage));
            if (sm) {
                const uint32_t r = base + (uint32_t)rank_in(sm);
                if (sus) {
                    const u32x4 h = L.suspended_header();
                    uint32_t* dst = &w->stage[r][0];
                    dst[0] = h.x;
                    dst[1] = h.y;
                    dst[2] = h.z;
                    dst[3] = h.w;
#pragma unroll
                    for (int i = 0; i < CtxShape<P>::kFWr; ++i) dst[4 + i] = L.frame[i];
                    w->stage_id[r] = L.tid | (L.lds ? 0u : kHeaderOnly);  // frame already in HBM
                }
                st.suspended += (uint32_t)popc(sm);
            }
            const uint32_t total = base + (uint32_t)popc(sm);
            if (total) {
                wave_sync();
                write_staged<P>(dev, L.rs, w, total);
                wave_sync();
                if (lane == 0) vst(&w->nstage, 0u);
            }
        }
        phase_mark(PH_STAGE);
        complete_stores();  // results, suspended records and child records of this round
        phase_mark(PH_WAIT);
        uint32_t ready = kNoParent, ready_q = 0;
        if (fin && fin_parent != kNoParent) {
            // the parent's join word: the last child also reads the resume queue
            const uint64_t old = atomic_add(reinterpret_cast<uint64_t*>(rec_ptr(dev, fin_parent) + kHdrJoin), fin_delta);
            const uint32_t left = fin_sum ? (uint32_t)old & kSumMaxCount : (uint32_t)old;
            if (left == 1u) {
                ready = fin_parent;
                ready_q = fin_sum ? ((uint32_t)old >> 12) & 0xFu : hdr_queue((uint32_t)(old >> 32));
            } else if (left == 0u) {
                GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_TASKWAIT, fin_tid, -1, 0, 0);
            }
        }
        if constexpr (Prof) {
            __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the join atomics have returned
            phase_mark(PH_JOIN);
        }
        if (ballot(ready != kNoParent)) {
            if (ready != kNoParent && ready >= dev.nrecs) complete_job<P>(dev, ready);
            if (ready != kNoParent && ready < dev.nrecs)
                queue_append<P, false>(dev, w, me, ready_q < (uint32_t)Q ? ready_q : 0u, ready, nullptr);
        }
        if (fin && fin_tid != 0u) free_id(dev, w, me, fin_tid);  // id 0 is the sink record

        // choose the busiest queue, keep its newest <= 64 tasks, publish the rest
        uint32_t kmax = 0, cmax = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint32_t c = uniform(vld(&w->tail[q]) - vld(&w->bottom[q]));
            if (c > cmax) {
                cmax = c;
                kmax = (uint32_t)q;
            }
        }
        const uint32_t keep_cap = ((int)kmax == WQ) ? 1u : (uint32_t)kWave;
        uint32_t keep = cmax < keep_cap ? cmax : keep_cap;
        uint32_t next_tid = 0, next_slot = 0;
        bool next_lds = false;
        if (keep) {
            const uint32_t t = uniform(vld(&w->tail[kmax]));
            const uint64_t vmask = uniform(vld(&w->valid[kmax]));
            const uint64_t lmask = uniform(vld(&w->inlds[kmax]));
            if ((uint32_t)lane < keep) {
                next_slot = (t - keep + (uint32_t)lane) & (kWave - 1);
                next_tid = vld(&w->mirror[kmax][next_slot]);
                next_lds = (lmask >> next_slot) & 1ull;
            }
            const uint64_t taken = rotl64(low_bits(keep), t - keep);
            if (lane == 0) {
                vst(&w->valid[kmax], vmask & ~taken);
                vst(&w->inlds[kmax], lmask & ~taken);
                vst(&w->tail[kmax], t - keep);
                vst(&w->cur_queue, kmax);
            }
        }
        // records pushed out of the window by the continuations appended above
        {
            const uint32_t ns = uniform(vld(&w->nstage));
            if (ns) {
                wave_sync();
                write_staged<P>(dev, L.rs, w, ns);
                wave_sync();
                if (lane == 0) vst(&w->nstage, 0u);
            }
        }
        // write out what is published: entries still in the 64-slot window (lane i takes
        // slot i) get their deque slot, and their record if it is held in LDS
        // Queues other than the kept one (several EPAQ queues) keep their 64-slot window
        // private for later rounds; only the entries pushed out of it are published.
        const __amdgpu_buffer_rsrc_t rs = L.rs;
        bool any_publish = false;
        uint32_t pub_to[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint32_t t = uniform(vld(&w->tail[q]));
            const uint32_t b = uniform(vld(&w->bottom[q]));
            const bool retain = Sched != kSchedGQ && Q > 1 && q < QL && (uint32_t)q != kmax && !dev.epaq_publish;
            pub_to[q] = !retain ? t : ((int32_t)(t - b) > kWave ? t - kWave : b);
            if (retain && pub_to[q] != b) any_publish = true;
            if (!retain && t != b) {
                any_publish = true;
                const uint64_t vmask = uniform(vld(&w->valid[q]));
                const uint64_t lmask = uniform(vld(&w->inlds[q]));
                const bool mine = (vmask >> lane) & 1ull;
                uint32_t oid = 0;
                if (mine) {
                    const uint32_t pos = t - kWave + (((uint32_t)lane - t) & (kWave - 1));
                    oid = vld(&w->mirror[q][lane]);
                    st32(dq_slot_ptr(dev, q, me, pos), oid);
                }
                if (q < QL) {
                    // records held in LDS: staged in record order, written together
                    const bool stg = mine && ((lmask >> lane) & 1ull);
                    const uint64_t sm = ballot(stg);
                    if (sm) {
                        if (stg) {
                            const uint32_t r = (uint32_t)rank_in(sm);
                            const uint32_t* d = &w->desc[q < QL ? q : 0][lane][0];
                            uint32_t* dst = &w->stage[r][0];
                            dst[0] = d[1];
                            dst[1] = d[2];
                            dst[2] = 0u;
                            dst[3] = d[0];
#pragma unroll
                            for (int i = 0; i < CtxShape<P>::kFWr; ++i) dst[4 + i] = (i < P::kFrameWords) ? d[3 + (i < P::kFrameWords ? i : 0)] : 0u;
                            w->stage_id[r] = oid;
                        }
                        st.published += (uint32_t)popc(sm);
                        wave_sync();
                        write_staged<P>(dev, rs, w, (uint32_t)popc(sm));
                        wave_sync();
                    }
                }
                if (lane == 0) {
                    vst(&w->valid[q], (uint64_t)0);
                    vst(&w->inlds[q], (uint64_t)0);
                }
            }
        }
        if (any_publish) {
            complete_stores();  // records and slots of everything published below
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const uint32_t b = uniform(vld(&w->bottom[q]));
                const uint32_t t = pub_to[q];
                if (t != b) {
                    if constexpr (Sched == kSchedGQ) {
                        gq_push<P>(dev, w, me, (uint32_t)q, b, t);
                    } else {
                        // top_cache stays a lower bound of the top; the capacity check re-reads it
                        if (lane == 0) atomic_add_noret(dq_word_ptr(dev, q, me), (uint64_t)(t - b));
                    }
                    vst(&w->bottom[q], t);
                }
            }
        }
        // a deque above half its capacity hands its oldest published tasks to the overflow ring
        if constexpr (Sched != kSchedGQ) {
            if (dev.ovf_cells != nullptr) {
#pragma unroll 1
                for (int q = 0; q < Q; ++q) {
                    const uint32_t b = uniform(vld(&w->bottom[q]));
                    if (b - uniform(vld(&w->top_cache[q])) > dev.cap / 2) {
                        const uint64_t r = ovf_spill_deque(ovf_view(dev), (uint32_t)q, dq_word_ptr(dev, q, me),
                                                           dq_slot_ptr(dev, q, me, 0), dev.cap);
                        if (lane == 0) {
                            vst(&w->top_cache[q], (uint32_t)r);
                            vst(&w->ring_spilled, vld(&w->ring_spilled) + (uint32_t)(r >> 32));
                        }
                    }
                }
            }
        }
        // A kept batch of fewer than 64 tasks (the queue then has no private entries left) is
        // topped up with the wave's own newest published tasks of the same queue, one CAS on
        // its deque word, so the next round runs up to 64 tasks instead of waiting until the
        // wave runs dry. Their records are in HBM and load at the round's start.
        if constexpr (Sched == kSchedWS) {
            if (keep < (uint32_t)kWave && (int)kmax != WQ) {
                uint32_t ptid = 0;
                const uint32_t k = try_pop<P>(dev, w, me, kmax, ptid, (uint32_t)kWave - keep);
                if (k) {
                    const int src = lane - (int)keep;
                    const uint32_t v = shfl(ptid, src < 0 ? 0 : src);
                    if ((uint32_t)lane >= keep && (uint32_t)lane < keep + k) {
                        next_tid = v;
                        next_lds = false;
                    }
                    keep += k;
                    st.pops += 1;
                    if (lane == 0) vst(&w->cur_queue, kmax);
                }
            }
        }
        phase_mark(PH_PUBLISH);
        n = keep;
        my_tid = next_tid;
        my_slot = next_slot;
        my_lds = next_lds;
        batch_q = (int)kmax;
        if (Prof && dev.prof_max && lane == 0) prof_stamp(dev.prof_working, dev.prof_counts, dev.prof_max, me, pw_idx, (int32_t)round_n);

        busy_iter += 1;
        if ((busy_iter & 63u) == 0) {
            int32_t a = 0;
            if (lane == 0) {
                a = ld32(dev.ctl + CTL_ABORT * 32);
                if (!a && (busy_iter & 1023u) == 0 && now_ticks() - t_begin > dev.timeout_ticks) {
                    GTAP_DEVICE_ERROR(dev, GTAP_ERROR_WATCHDOG, -1, -1, (int)((now_ticks() - t_begin) / 100000u),
                                      (int)(dev.timeout_ticks / 100000u));
                    a = 1;
                }
            }
            a = shfl(a, 0);
            if (a) break;
        }
    }

    if (lane == 0) {
        uint64_t* s = dev.stats + (size_t)me * ST_COUNT;
        s[ST_TASKS] = st.tasks;
        s[ST_SPAWNS] = vld(&w->spawns);
        s[ST_ROUNDS] = st.rounds;
        s[ST_POPS] = st.pops;
        s[ST_STEALS] = st.steals;
        s[ST_STEAL_ATTEMPTS] = st.attempts;
        s[ST_JOBS] = st.jobs;
        s[ST_REFILLS] = vld(&w->refills);
        s[ST_PUBLISHED] = st.published;
        s[ST_OVERFLOWED] = vld(&w->overflowed);
        s[ST_SUSPENDED] = st.suspended;
        s[ST_START_TICKS] = t_begin;
        s[ST_END_TICKS] = now_ticks();
        s[ST_RING_SPILLED] = vld(&w->ring_spilled);
        s[ST_RING_TAKEN] = st.ring_taken;
        if constexpr (Prof) {
            if (dev.prof_phase != nullptr) {
                uint64_t* ph = dev.prof_phase + (size_t)me * PH_COUNT;
#pragma unroll
                for (int k = 0; k < PH_ROUNDS; ++k) ph[k] = phase_acc[k];
                ph[PH_ROUNDS] = st.rounds;
            }
        }
    }
}

// Launch shape: 256-thread workgroups (4 waves), one WaveCtx per wave in LDS.
template <class P, class = void>
struct WaveSharedOf {
    struct type {
        uint32_t unused;
    };
};
template <class P>
struct WaveSharedOf<P, decltype((void)sizeof(typename P::WaveShared))> {
    using type = typename P::WaveShared;
};

template <class P, int Sched = kSchedWS, bool Prof = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WavesPerSimd<P>::value)))
void thread_kernel(GtapDev dev, typename P::Args args) {
    __shared__ WaveCtxOf<P> ctxs[4];
    __shared__ typename WaveSharedOf<P>::type wsh[4];
    thread_scheduler<P, Sched, Prof>(dev, args, &ctxs[threadIdx.x / kWave], &wsh[threadIdx.x / kWave]);
}

}  // namespace gtap
