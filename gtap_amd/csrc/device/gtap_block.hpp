// gtap_block.hpp -- block-cooperative persistent scheduler (one task per 256-thread
// workgroup at a time), the counterpart of the reference's block mode
// (runtime/detail/block/gtap_runtime_block.cuh:326-777, gtap_block_core.cuh).
//
//  * Worker = one workgroup of 256 threads (4 waves). Every thread enters the task
//    function; task bodies are written data-parallel over threadIdx.x. taskwait, return and
//    finish are workgroup-collective (they contain barriers), spawn may be called by any
//    subset of threads.
//  * Same task records, id pool and deque words as the thread-mode scheduler. A workgroup
//    owns one deque; it publishes with one atomic add, pops one task from the bottom and
//    thieves take one task from the top, each with one CAS on the (top | bottom) word.
//  * Work first (reference: push, gtap_runtime_block.cuh:400-438): after a task the
//    workgroup continues with the parent it just made ready, else with its newest child,
//    and publishes the other children.
//  * Thread 0 runs the scheduler; the others wait at a barrier.
#pragma once

#include "gtap_thread.hpp"

namespace gtap {

constexpr uint32_t kBlockThreads = 256;
constexpr uint32_t kBlockMaxSpawn = 128;   // ids kept ready in LDS for one task invocation's spawns

// Scheduler loop state of thread 0, kept in LDS rather than in registers: it stays live
// across every task body, where registers are scarce.
struct BlockLoop {
    uint64_t t_begin;
    uint64_t host_poll_last;
    uint32_t rng;
    uint32_t idle_iter;
    uint32_t ph_idx, pw_idx;  // timeline profiling
    uint32_t seen_next;       // last value of the admission counter read
    bool idle_registered;
    bool had_task;
    bool static_job;          // this workgroup still has to start job `me` (handed out by index)
    uint32_t tasks, pops, steals, attempts, jobs;
};

struct BlockCtx {
    BlockLoop lp;
    // scheduler state (thread 0)
    uint32_t cur;          // task to run next (0 = none)
    uint32_t exit;
    uint32_t tail;         // private tail of the deque (absolute position)
    uint32_t bottom;       // published bottom
    uint32_t top_cache;
    uint32_t free_top;     // ids in free_ids
    uint32_t priv_top;     // ids in this workgroup's HBM overflow stack
    uint32_t child_count;  // joined children spawned by the current invocation
    uint32_t taken;        // ids taken from free_ids by the current invocation
    uint32_t priv_taken;   // ids taken from the HBM stack by the current invocation (LDS ran out)
    uint32_t spawns;
    uint32_t refills;
    // header of the current task (written by thread 0, read by all)
    uint32_t tid, func, state, parent, res_off, outcome;
    uint32_t frame[16];
    uint32_t free_ids[kFreeCap];
};

__device__ __forceinline__ uint32_t blk_thread() { return threadIdx.x; }

// Block-collective refill of the LDS id stack: one chunk from the workgroup's HBM stack, else
// one ring cell, else `want` chunks of never-used ids with one fetch-add. The first wave
// probes every ring at once (one lane per ring); thread 0 claims; all copy.
template <int DUMMY = 0>
__device__ __forceinline__ void block_refill_chunk(const GtapDev& dev, BlockCtx* c, uint32_t me, uint32_t want) {
    __shared__ uint32_t s_kind, s_pos, s_shard, s_fresh, s_nch;
    const uint32_t t = blk_thread();
    const uint32_t top = c->free_top;
    const uint32_t ptop = c->priv_top;
    if (ptop >= kChunk) {
        const uint32_t* src = dev.priv_ids + (size_t)me * kPrivCap + (ptop - kChunk);
        if (t < kChunk) c->free_ids[top + t] = src[t];
        __syncthreads();
        if (t == 0) {
            c->priv_top = ptop - kChunk;
            c->free_top = top + kChunk;
        }
        __syncthreads();
        return;
    }
    if (t < (uint32_t)kWave) {
        uint32_t kind = 0, pos = 0, shard = 0, fresh = 0, nch = 0;
        bool fresh_left = true;
        for (int attempt = 0; attempt < 64; ++attempt) {
            const bool seen = ring_probe(dev, me, (uint32_t)kWave, t, shard, pos);
            if (seen) {
                int ok = 0;
                if (t == 0) {
                    uint32_t expected = pos;
                    ok = atomic_cas(pool_deq_ptr(dev, shard), expected, pos + 1) ? 1 : 0;
                }
                if (shfl((int32_t)ok, 0)) {
                    kind = 1;
                    nch = 1;
                    break;
                }
            }
            if (fresh_left) {
                uint32_t f = 0, got = 0;
                if (t == 0) {
                    f = atomic_add(pool_fresh_ptr(dev), want * kChunk);
                    const uint32_t left = f < dev.nrecs ? (dev.nrecs - f) / kChunk : 0u;
                    got = left < want ? left : want;
                }
                fresh = shfl(f, 0);
                nch = shfl(got, 0);
                if (nch) {
                    kind = 2;
                    break;
                }
                fresh_left = false;
            }
            if (!seen) break;  // every ring empty and no never-used ids left
        }
        if (t == 0) {
            s_kind = kind;
            s_pos = pos;
            s_shard = shard;
            s_fresh = fresh;
            s_nch = kind ? nch : 1u;
            c->refills += 1;
        }
    }
    __syncthreads();
    const uint32_t kind = s_kind;
    const uint32_t nids = s_nch * kChunk;
    for (uint32_t i = t; i < nids; i += kBlockThreads) {
        uint32_t id = 0;
        if (kind == 1) id = ld32(pool_cell_ptr(dev, s_shard, s_pos) + i);
        else if (kind == 2) id = s_fresh + i;
        c->free_ids[top + i] = id;
    }
    complete_stores();
    __syncthreads();
    if (t == 0) {
        if (kind == 1) st32(pool_seq_ptr(dev, s_shard, s_pos), s_pos + dev.ncells);
        if (kind == 0) GTAP_DEVICE_ERROR(dev, GTAP_ERROR_TASK_ID_POOL_LOW_HEADROOM, -1, -1, (int)s_fresh, (int)dev.nrecs);
        c->free_top = top + nids;
    }
    __syncthreads();
}

// Block-collective spill of the 64 most recently freed ids.
template <int DUMMY = 0>
__device__ __forceinline__ void block_spill_chunk(const GtapDev& dev, BlockCtx* c, uint32_t me) {
    __shared__ uint32_t s_ok, s_pos;
    const uint32_t t = blk_thread();
    const uint32_t top = c->free_top;
    const uint32_t ptop = c->priv_top;
    if (ptop + kChunk <= dev.priv_cap) {
        uint32_t* dst = dev.priv_ids + (size_t)me * kPrivCap + ptop;
        if (t < kChunk) dst[t] = c->free_ids[top - kChunk + t];
        __syncthreads();
        if (t == 0) {
            c->priv_top = ptop + kChunk;
            c->free_top = top - kChunk;
        }
        __syncthreads();
        return;
    }
    const uint32_t shard = me % dev.nshards;
    if (t == 0) {
        uint32_t pos = 0;
        s_ok = ring_put_claim(dev, shard, pos) ? 1u : 0u;
        s_pos = pos;
    }
    __syncthreads();
    if (s_ok && t < kChunk) st32(pool_cell_ptr(dev, shard, s_pos) + t, c->free_ids[top - kChunk + t]);
    complete_stores();
    __syncthreads();
    if (t == 0) {
        if (s_ok) st32(pool_seq_ptr(dev, shard, s_pos), s_pos + 1);
        else GTAP_DEVICE_ERROR(dev, GTAP_ERROR_TASK_ID_POOL_SLOT_BUSY, -1, (int)shard, 0, (int)dev.ncells);
        c->free_top = top - kChunk;
    }
    __syncthreads();
}

// Task context seen by every thread of the workgroup.
template <class P>
struct BlockTask {
    const GtapDev* dev;
    BlockCtx* c;
    const typename P::Args* args;
    typename P::Shared* sh;    // program scratch in LDS
    uint32_t me;               // workgroup index
    uint32_t tid, func, state, parent, res_off;
    uint8_t* rec;

    __device__ __forceinline__ uint32_t get_state() const { return state; }
    template <typename T>
    __device__ __forceinline__ T arg(int byte_off) const {
        if constexpr (sizeof(T) == 4) {
            return __builtin_bit_cast(T, c->frame[byte_off / 4]);
        } else {
            const uint64_t v = (uint64_t)c->frame[byte_off / 4] | ((uint64_t)c->frame[byte_off / 4 + 1] << 32);
            return __builtin_bit_cast(T, v);
        }
    }
    template <typename T>
    __device__ __forceinline__ T frame_ld(int byte_off) const { return ldT<T>(rec + kHdrBytes + byte_off); }
    template <typename T>
    __device__ __forceinline__ void frame_st(int byte_off, T v) { stT<T>(rec + kHdrBytes + byte_off, v); }

    // Any thread may spawn. Returns the child id.
    template <typename A>
    __device__ __forceinline__ uint32_t spawn(uint32_t f, int res_off_in_frame, const A& a) {
        return spawn_impl(f, tid, kHdrBytes + (uint32_t)res_off_in_frame, a);
    }
    template <typename A>
    __device__ __forceinline__ uint32_t spawn_detached(uint32_t f, const A& a) {
        return spawn_impl(f, kNoParent, 0u, a);
    }
    template <typename A>
    __device__ __forceinline__ uint32_t spawn_impl(uint32_t f, uint32_t par, uint32_t roff, const A& a) {
        static_assert(sizeof(A) % 8 == 0 && sizeof(A) / 4 <= P::kFrameWords, "argument block");
        // free_top and priv_top do not change while a task runs; spawns claim ids below them:
        // the LDS stack first, then this workgroup's HBM stack, then never-used ids (a task
        // may spawn any number of children, as a vertex task of a block-mode BFS does)
        const uint32_t k = atomicAdd(&c->taken, 1u);
        const uint32_t avail = c->free_top;
        uint32_t id = 0;
        if (k < avail) {
            id = c->free_ids[avail - 1 - k];
        } else {
            const uint32_t k2 = atomicAdd(&c->priv_taken, 1u);
            const uint32_t ptop = c->priv_top;
            if (k2 < ptop) {
                id = dev->priv_ids[(size_t)me * kPrivCap + (ptop - 1 - k2)];
            } else {
                const uint32_t f = atomic_add(pool_fresh_ptr(*dev), 1u);
                if (f < dev->nrecs) id = f;
                else GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_GENERATED_TASK_ID_BUFFER_OVERFLOW, tid, -1, (int)k, (int)avail);
            }
        }
        uint8_t* crec = rec_ptr(*dev, id);
        st64(reinterpret_cast<uint64_t*>(crec), hdr_lo(par, roff));
        st64(reinterpret_cast<uint64_t*>(crec + kHdrJoin), hdr_hi(0u, hdr_word0(f, 0, 0)));
        const uint64_t* src = reinterpret_cast<const uint64_t*>(&a);
        uint64_t* dst = reinterpret_cast<uint64_t*>(crec + kHdrBytes);
#pragma unroll
        for (int i = 0; i < (int)(sizeof(A) / 8); ++i) st64(dst + i, src[i]);
        if (id != 0u) {
            const uint32_t pos = atomicAdd(&c->tail, 1u);
            if (pos + 1u - c->top_cache > dev->cap - kMargin) {
                GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_QUEUE_OVERFLOW, id, 0, pos + 1u - c->top_cache, dev->cap - kMargin);
            } else {
                st32(dq_slot_ptr(*dev, 0, me, pos), id);
            }
            if (par != kNoParent) atomicAdd(&c->child_count, 1u);
            atomicAdd(&c->spawns, 1u);
        }
        return id;
    }

    // Workgroup-collective taskwait: true -> return now, the task resumes at next_state.
    __device__ __forceinline__ bool taskwait(uint32_t next_state) {
        complete_stores();
        __syncthreads();
        const uint32_t cc = c->child_count;
        if (cc == 0) {
            state = next_state;
            __syncthreads();
            return false;
        }
        if (threadIdx.x == 0) {
            st64(reinterpret_cast<uint64_t*>(rec + kHdrJoin), hdr_hi(cc, hdr_word0(func, next_state, 0)));
            complete_stores();
            c->outcome = OUT_SUSPENDED;
        }
        __syncthreads();
        return true;
    }
    // Workgroup-collective return; thread 0's value is stored.
    template <typename T>
    __device__ __forceinline__ void ret(T v) {
        if (threadIdx.x == 0 && parent != kNoParent) stT<T>(rec_ptr(*dev, parent) + res_off, v);
        finish();
    }
    __device__ __forceinline__ void finish() {
        complete_stores();
        __syncthreads();
        if (threadIdx.x == 0) {
            if (c->child_count != 0) GTAP_DEVICE_ERROR(*dev, GTAP_ERROR_INVALID_TASKWAIT, tid, -1, c->child_count, 0);
            c->outcome = OUT_FINISHED;
        }
        __syncthreads();
    }
};

// Thread 0 only: pop the newest published task of this workgroup's deque.
// A failed CAS means a thief took tasks, so the loop ends with a task or an empty deque: a
// workgroup that counts itself idle has nothing left in its own deque.
__device__ __forceinline__ uint32_t block_pop(const GtapDev& dev, BlockCtx* c, uint32_t me) {
    uint64_t* wp = dq_word_ptr(dev, 0, me);
    while (true) {
        const uint64_t wv = ld64(wp);
        const uint32_t t = (uint32_t)(wv >> 32), b = (uint32_t)wv;
        if ((int32_t)(b - t) <= 0) return 0;
        uint64_t expected = wv;
        if (atomic_cas(wp, expected, ((uint64_t)t << 32) | (uint64_t)(b - 1))) {
            const uint32_t id = ld32(dq_slot_ptr(dev, 0, me, b - 1));
            c->bottom = b - 1;
            c->tail = b - 1;
            c->top_cache = t;
            return id;
        }
    }
}

// Thread 0 only: steal the oldest half (at most kBlockStealMax, and at most the configured
// steal batch) of a random victim's published tasks. The oldest one runs here; the others
// are pushed onto this workgroup's own deque and published at once, so a wide fan-out
// reaches idle workgroups in a logarithmic number of steal rounds instead of one steal per
// task from the spawner. Claimed slots are read after the CAS; the owner never writes
// within kMargin of the top it last saw, so they cannot be overwritten first.
constexpr uint32_t kBlockStealMax = 64;
__device__ __forceinline__ uint32_t block_steal(const GtapDev& dev, uint32_t me, BlockCtx* c) {
    BlockLoop& lp = c->lp;
    if (dev.nworkers < 2) return 0;
    for (int probe = 0; probe < 4; ++probe) {
        uint32_t rng = lp.rng;
        uint32_t victim = xorshift32(rng) % (dev.nworkers - 1);
        lp.rng = rng;
        if (victim >= me) victim += 1;
        uint64_t* wp = dq_word_ptr(dev, 0, victim);
        const uint64_t wv = ld64(wp);
        const uint32_t t = (uint32_t)(wv >> 32), b = (uint32_t)wv;
        const int32_t sz = (int32_t)(b - t);
        if (sz <= 0) continue;
        uint32_t k = ((uint32_t)sz + 1u) / 2u;
        const uint32_t cap = dev.steal_batch < kBlockStealMax ? dev.steal_batch : kBlockStealMax;
        if (k > cap) k = cap;
        if (k == 0) k = 1;
        if (dev.term_mode == GTAP_TERM_QUIESCE && lp.idle_registered) {
            atomic_add(dev.ctl + CTL_IDLE * 32, -1);
            lp.idle_registered = false;
        }
        uint64_t expected = wv;
        lp.attempts += 1;
        if (!atomic_cas(wp, expected, wv + ((uint64_t)k << 32))) return 0;
        lp.steals += 1;
        const uint32_t first = ld32(dq_slot_ptr(dev, 0, victim, t));
        if (k > 1) {
            uint32_t pos = c->tail;
            for (uint32_t i0 = 1; i0 < k; i0 += 16) {  // 16 slot loads in flight at a time
                uint32_t ids[16];
#pragma unroll
                for (uint32_t i = 0; i < 16; ++i)
                    if (i0 + i < k) ids[i] = ld32(dq_slot_ptr(dev, 0, victim, t + i0 + i));
#pragma unroll
                for (uint32_t i = 0; i < 16; ++i) {
                    if (i0 + i < k) {
                        if (pos + 1u - c->top_cache > dev.cap - kMargin) {
                            GTAP_DEVICE_ERROR(dev, GTAP_ERROR_QUEUE_OVERFLOW, ids[i], 0, pos + 1u - c->top_cache,
                                              dev.cap - kMargin);
                        } else {
                            st32(dq_slot_ptr(dev, 0, me, pos), ids[i]);
                            pos += 1;
                        }
                    }
                }
            }
            c->tail = pos;
            complete_stores();  // the slots are written before they are published
            const uint64_t old = atomic_add(dq_word_ptr(dev, 0, me), (uint64_t)(c->tail - c->bottom));
            c->top_cache = (uint32_t)(old >> 32);
            c->bottom = c->tail;
        }
        return first;
    }
    return 0;
}

// Thread 0 only: job j of the table becomes a root task (anchor record with join = 1, root
// record with the job's arguments) and is marked running. Returns the root id.
template <class P>
__device__ __forceinline__ uint32_t block_start_job(const GtapDev& dev, BlockCtx* c, uint32_t j) {
    const uint32_t slot = j % dev.job_ring;
    gtap_job* job = dev.jobs + slot;
    const uint32_t func = ld32_sys(&job->func);
    const uint32_t anchor = dev.nrecs + slot;
    c->free_top -= 1;  // free_top >= kBlockMaxSpawn here (refilled before acquisition)
    const uint32_t id = c->free_ids[c->free_top];
    uint8_t* arec = rec_ptr(dev, anchor);
    st64(reinterpret_cast<uint64_t*>(arec), hdr_lo(kNoParent, 0u));
    st64(reinterpret_cast<uint64_t*>(arec + kHdrJoin), hdr_hi(1u, hdr_word0(kAnchorFunc, 0, 0)));
    st64(reinterpret_cast<uint64_t*>(arec + kHdrBytes), 0ull);
    uint8_t* rrec = rec_ptr(dev, id);
    st64(reinterpret_cast<uint64_t*>(rrec), hdr_lo(anchor, kHdrBytes));
    st64(reinterpret_cast<uint64_t*>(rrec + kHdrJoin), hdr_hi(0u, hdr_word0(func, 0, 0)));
    const uint64_t* ja = reinterpret_cast<const uint64_t*>(job->args);
    uint64_t* ra = reinterpret_cast<uint64_t*>(rrec + kHdrBytes);
    constexpr int kArgWords = (P::kFrameWords / 2) < 8 ? (P::kFrameWords / 2) : 8;
#pragma unroll
    for (int i = 0; i < kArgWords; ++i) st64(ra + i, ld64_sys(ja + i));
    st64_sys(&job->t_start, now_ticks());
    st32_sys(&job->status, (uint32_t)GTAP_JOB_RUNNING);
    complete_stores();
    return id;
}

// Thread 0 only: admit one job from the job table as a root task. `seen_next` is the last
// value of the admission counter this workgroup read; the counter only grows, so once it
// was seen at the number of jobs available its line is not read again.
template <class P>
__device__ __forceinline__ uint32_t block_admit(const GtapDev& dev, BlockCtx* c, bool& idle_registered,
                                                uint32_t& seen_next) {
    uint32_t avail = dev.njobs;
    if (dev.term_mode == GTAP_TERM_SERVE) avail = (uint32_t)ld32(dev.ctl + CTL_HOST_AVAIL * 32);
    if (seen_next >= avail) return 0;
    int32_t* nextp = dev.ctl + CTL_JOB_NEXT * 32;
    const uint32_t next = (uint32_t)ld32(nextp);
    seen_next = next;
    if (next >= avail) return 0;
    const uint32_t live = (uint32_t)ld32(dev.ctl + CTL_LIVE_JOBS * 32);
    if (live >= dev.max_live_jobs) return 0;
    if (c->free_top == 0) return 0;
    if (dev.term_mode == GTAP_TERM_QUIESCE && idle_registered) {
        atomic_add(dev.ctl + CTL_IDLE * 32, -1);
        idle_registered = false;
    }
    // live count first, admission counter second (undone if the claim fails), so a stop rule
    // reading them in the other order never misses an admitted job
    atomic_add(dev.ctl + CTL_LIVE_JOBS * 32, 1);
    complete_stores();
    int32_t expected = (int32_t)next;
    if (!atomic_cas(nextp, expected, (int32_t)(next + 1))) {
        atomic_add_noret(dev.ctl + CTL_LIVE_JOBS * 32, -1);
        return 0;
    }
    return block_start_job<P>(dev, c, next);
}

template <class P>
__device__ void block_scheduler(const GtapDev& dev, const typename P::Args& args, BlockCtx* c,
                                typename P::Shared* sh) {
    const uint32_t t = blk_thread();
    const uint32_t me = blockIdx.x;
    BlockLoop& lp = c->lp;
    if (t == 0) {
        lp.t_begin = now_ticks();
        // a workgroup that starts a job by index was counted as started by the reset
        if (dev.term_mode == GTAP_TERM_QUIESCE && me >= dev.static_jobs) atomic_add(dev.ctl + CTL_STARTED * 32, 1);
        lp.host_poll_last = 0;
        lp.rng = 0x9E3779B9u ^ (me * 0x85EBCA6Bu) ^ (uint32_t)lp.t_begin;
        if (lp.rng == 0) lp.rng = 1;
        lp.idle_iter = 0;
        lp.ph_idx = lp.pw_idx = 0;
        lp.idle_registered = false;
        lp.had_task = false;
        lp.seen_next = dev.jobs != nullptr ? dev.static_jobs : 0u;
        lp.static_job = dev.jobs != nullptr && me < dev.static_jobs;
        lp.tasks = lp.pops = lp.steals = lp.attempts = lp.jobs = 0;
        c->cur = 0;
        c->exit = 0;
        c->tail = c->bottom = c->top_cache = 0;
        c->free_top = 0;
        c->priv_top = 0;
        c->child_count = 0;
        c->taken = 0;
        c->priv_taken = 0;
        c->spawns = 0;
        c->refills = 0;
    }
    __syncthreads();
    // the workgroup's first ids are a fixed range (the shared fresh-id counter starts above
    // every worker's range)
    if (dev.static_ids) {
        for (uint32_t i = t; i < dev.static_ids; i += kBlockThreads) c->free_ids[i] = 64u + me * dev.static_ids + i;
        __syncthreads();
        if (t == 0) c->free_top = dev.static_ids;
        __syncthreads();
    }

    BlockTask<P> T;
    T.dev = &dev;
    T.c = c;
    T.args = &args;
    T.sh = sh;
    T.me = me;

    const bool host_poller = dev.term_mode == GTAP_TERM_SERVE && me == 0 && t == 0;

    while (true) {
        if (host_poller) host_poll(dev, lp.host_poll_last);
        // keep enough ids in LDS for one invocation's spawns (collective)
        while (c->free_top < kBlockMaxSpawn)
            block_refill_chunk(dev, c, me, (kBlockMaxSpawn - c->free_top + kChunk - 1) / kChunk);
        if (t == 0 && c->cur == 0) {
            uint32_t id = 0;
            if (lp.static_job) {
                lp.static_job = false;
                id = block_start_job<P>(dev, c, me);
                lp.jobs += 1;
            }
            if (!id) {
                id = block_pop(dev, c, me);
                if (id) lp.pops += 1;
            }
            if (!id && dev.jobs != nullptr) {
                id = block_admit<P>(dev, c, lp.idle_registered, lp.seen_next);
                if (id) lp.jobs += 1;
            }
            if (!id) id = block_steal(dev, me, c);
            if (!id) {
                if (lp.had_task) {
                    lp.had_task = false;
                    if (dev.prof_max) prof_stamp(dev.prof_having, nullptr, dev.prof_max, me, lp.ph_idx, 0);
                }
                if (dev.term_mode == GTAP_TERM_QUIESCE && !lp.idle_registered) {
                    lp.idle_registered = true;
                    const int32_t cnt = atomic_add(dev.ctl + CTL_IDLE * 32, 1) + 1;
                    // every started workgroup is idle (one not yet resident holds no work)
                    if (cnt == ld32(dev.ctl + CTL_STARTED * 32) &&
                        (uint32_t)ld32(dev.ctl + CTL_JOB_NEXT * 32) >= dev.njobs)
                        st32(dev.ctl + CTL_DONE * 32, 1);
                }
                const uint32_t idle_iter = ++lp.idle_iter;
                int32_t flags = ld32(dev.ctl + CTL_DONE * 32);  // an abort sets it too
                if ((idle_iter & 63u) == 0 && now_ticks() - lp.t_begin > dev.timeout_ticks) {
                    GTAP_DEVICE_ERROR(dev, GTAP_ERROR_WATCHDOG, -1, -1, (int)((now_ticks() - lp.t_begin) / 100000u),
                                      (int)(dev.timeout_ticks / 100000u));
                    flags |= 2;
                }
                if (flags) c->exit = 1;
                else {
                    const uint32_t lvl = idle_iter < 8 ? 0u : (idle_iter < 64 ? 2u : (idle_iter < 256 ? 4u : 6u));
                    backoff_sleep(lvl < dev.idle_level ? lvl : dev.idle_level);
                }
            } else {
                lp.idle_iter = 0;
                c->cur = id;
            }
            if (c->cur) {
                // load the task header and frame for everyone
                const uint32_t cur = c->cur;
                if (cur >= dev.nrecs) {
                    GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_FUNCTION, cur, -1, cur, dev.nrecs);
                    c->cur = 0;
                    c->exit = 1;
                }
            }
        }
        __syncthreads();
        if (c->exit) break;
        if (c->cur == 0) continue;
        if (t == 0) {
            uint8_t* r = rec_ptr(dev, c->cur);
            const uint64_t h0 = ld64(reinterpret_cast<const uint64_t*>(r));
            const uint64_t h1 = ld64(reinterpret_cast<const uint64_t*>(r + kHdrJoin));
            c->tid = c->cur;
            c->func = hdr_func((uint32_t)(h1 >> 32));
            c->state = hdr_state((uint32_t)(h1 >> 32));
            c->parent = (uint32_t)h0;
            c->res_off = (uint32_t)(h0 >> 32);
            c->outcome = OUT_RUNNING;
            c->child_count = 0;
            const uint64_t* fp = reinterpret_cast<const uint64_t*>(r + kHdrBytes);
#pragma unroll
            for (int i = 0; i < P::kFrameWords / 2; ++i) {
                const uint64_t v = ld64(fp + i);
                c->frame[2 * i] = (uint32_t)v;
                c->frame[2 * i + 1] = (uint32_t)(v >> 32);
            }
        }
        __syncthreads();
        T.tid = c->tid;
        T.func = c->func;
        T.state = c->state;
        T.parent = c->parent;
        T.res_off = c->res_off;
        T.rec = rec_ptr(dev, T.tid);
        if (t == 0) {
            lp.tasks += 1;
            if (dev.prof_max) {
                if (!lp.had_task) prof_stamp(dev.prof_having, nullptr, dev.prof_max, me, lp.ph_idx, 0);
                prof_stamp(dev.prof_working, dev.prof_counts, dev.prof_max, me, lp.pw_idx, 1);
            }
            lp.had_task = true;
        }
        if (T.func == 0 || T.func > P::kMaxFunc) {
            if (t == 0) GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_FUNCTION, T.tid, 0, T.func, P::kMaxFunc);
            T.finish();
        } else {
            P::run(T);
        }
        __syncthreads();
        // ---- epilogue (thread 0): join, release, choose the next task, publish the rest
        if (t == 0) {
            if (c->outcome == OUT_RUNNING) {
                GTAP_DEVICE_ERROR(dev, GTAP_ERROR_INVALID_TASKWAIT, T.tid, 0, T.state, 0);
                c->outcome = OUT_FINISHED;
            }
            uint32_t next = 0;
            {
                const uint32_t used = c->taken < c->free_top ? c->taken : c->free_top;
                c->free_top -= used;
                c->taken = 0;
                const uint32_t pused = c->priv_taken < c->priv_top ? c->priv_taken : c->priv_top;
                c->priv_top -= pused;
                c->priv_taken = 0;
            }
            if (c->outcome == OUT_FINISHED) {
                const uint32_t par = c->parent;
                if (par != kNoParent && par < dev.nrecs + dev.nanchors) {
                    const uint32_t old = atomic_sub(reinterpret_cast<uint32_t*>(rec_ptr(dev, par) + kHdrJoin), 1u);
                    if (old == 1u) {
                        if (par >= dev.nrecs) complete_job<P>(dev, par);
                        else next = par;  // resume the parent right here
                    }
                }
                if (T.tid != 0u) {
                    c->free_ids[c->free_top] = T.tid;  // room is ensured by the spill below
                    c->free_top += 1;
                }
            }
            if (next == 0 && c->tail != c->bottom) {
                c->tail -= 1;  // newest child runs next
                next = ld32(dq_slot_ptr(dev, 0, me, c->tail));
            }
            if (c->tail != c->bottom) {
                const uint64_t old = atomic_add(dq_word_ptr(dev, 0, me), (uint64_t)(c->tail - c->bottom));
                c->top_cache = (uint32_t)(old >> 32);
                c->bottom = c->tail;
            }
            c->cur = next;
            if (dev.prof_max) prof_stamp(dev.prof_working, dev.prof_counts, dev.prof_max, me, lp.pw_idx, 1);
            if ((lp.tasks & 63u) == 0 && ld32(dev.ctl + CTL_ABORT * 32)) c->exit = 1;
        }
        __syncthreads();
        if (c->exit) break;
        while (c->free_top > kFreeCap - kChunk) block_spill_chunk(dev, c, me);
    }

    if (t == 0) {
        uint64_t* s = dev.stats + (size_t)me * ST_COUNT;
        s[ST_TASKS] = lp.tasks;
        s[ST_SPAWNS] = c->spawns;
        s[ST_ROUNDS] = lp.tasks;
        s[ST_POPS] = lp.pops;
        s[ST_STEALS] = lp.steals;
        s[ST_STEAL_ATTEMPTS] = lp.attempts;
        s[ST_JOBS] = lp.jobs;
        s[ST_REFILLS] = c->refills;
        s[ST_START_TICKS] = lp.t_begin;
        s[ST_END_TICKS] = now_ticks();
    }
}

template <class P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WavesPerSimd<P>::value)))
void block_kernel(GtapDev dev, typename P::Args args) {
    __shared__ BlockCtx ctx;
    __shared__ typename P::Shared sh;
    block_scheduler<P>(dev, args, &ctx, &sh);
}

}  // namespace gtap
