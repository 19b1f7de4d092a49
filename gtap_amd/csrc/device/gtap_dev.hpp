// gtap_dev.hpp -- device-side runtime descriptor and CDNA4 primitives shared by the
// thread-mode and block-mode schedulers.
//
// Coherence model (gfx950, 8 XCDs with one L2 each):
//  * Every word that one worker writes and another worker reads (task records, deque
//    slots, id-pool cells, job entries) is accessed with agent-scope relaxed atomics,
//    which compile to `sc1` loads/stores. They are coherent across XCDs without any
//    cache maintenance.
//  * Ordering is "complete, then publish": a wave issues its sc1 stores, waits for them
//    with `s_waitcnt vmcnt(0)`, and only then performs the atomic that makes them
//    reachable (deque publish, join decrement, pool-cell release). A reader only issues
//    its sc1 loads after the atomic that handed it the work has returned.
//  * Agent-scope acquire/release fences are never used: each one is a whole-L2
//    writeback or invalidate, which costs tens of microseconds when thousands of waves
//    issue them (measured: 69 us per release fence at 8192 waves).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../gtap_types.h"

namespace gtap {

constexpr int kWave = 64;

// Control words, each on its own 128-byte line (index * 32 ints).
enum CtlIndex : int {
    CTL_DONE = 0,       // nonzero: all waves leave the scheduler loop (1 finished, bit 1 aborted)
    CTL_ABORT = 1,      // set on a device error or watchdog expiry
    CTL_IDLE = 2,       // idle-worker count (quiescence termination)
    CTL_ERROR = 3,      // first error code (CAS from 0)
    CTL_JOB_NEXT = 4,   // next job index to admit
    CTL_JOBS_DONE = 5,  // completed jobs
    CTL_LIVE_JOBS = 6,  // admitted and not yet completed
    CTL_HOST_AVAIL = 7,   // serve mode: device mirror of HCTL_JOB_AVAIL
    CTL_HOST_STOP = 8,    // serve mode: device mirror of HCTL_STOP
    CTL_STARTED = 9,      // workers that have entered the scheduler loop, plus the workers
                          // that start a job by index, counted from the reset (quiescence
                          // counts these, not the launch's worker count: a worker the hardware
                          // has not made resident yet and that owns no job holds no work and
                          // must not be waited for; one that owns a job must)
    CTL_PROGRAM = 10,     // a counter a program may use (cleared with the others per launch)
    CTL_ARMED = 11,       // 1 after gtap_initialize / gtap_reset: one host-launched entry kernel
                          // may run on this state (it clears the word when it is done)
    CTL_ADMIT_HOLDS = 12, // served: admission attempts the reservation ledger held back
    CTL_COUNT = 13
};
// Served admission ledger: the 8-byte word at CTL_JOB_NEXT holds the admission counter in its
// low half and, in its high half, the ids (in 64-id units) reserved by the admitted jobs that
// have not completed. One CAS admits jobs and reserves their ids together.

// Host-visible control (pinned host memory, read with system-scope loads).
enum HostCtlIndex : int {
    HCTL_JOB_AVAIL = 0,  // jobs the host has made available (serve mode)
    HCTL_STOP = 1,       // host asks the served kernel to drain and exit
    HCTL_COUNT = 2
};

// Per-wave statistics slots (uint64 each, 16 per wave).
enum StatIndex : int {
    ST_TASKS = 0, ST_SPAWNS, ST_ROUNDS, ST_POPS, ST_STEALS, ST_STEAL_ATTEMPTS, ST_JOBS, ST_REFILLS,
    ST_PUBLISHED, ST_OVERFLOWED, ST_SUSPENDED,
    ST_START_TICKS,  // wall-clock tick at which the worker entered the scheduler loop
    ST_END_TICKS,    // wall-clock tick at which it left the loop
    ST_RING_SPILLED, // tasks this worker moved from a full deque to the overflow ring
    ST_RING_TAKEN,   // tasks this worker took from the overflow ring
    ST_COUNT = 16
};

// Everything a persistent launch needs, passed by value as a kernel argument.
struct GtapDev {
    uint8_t* recs;             // task records, `stride` bytes each
    uint32_t* priv_ids;        // per-wave overflow id stacks, [worker][8192]
    uint32_t* pool_ids;        // id rings, [shard][cell][64]
    uint32_t* pool_seq;        // per-cell sequence numbers, [shard][cell]
    uint64_t* pool_ctl;        // [32s] enqueue pos, [32s+16] dequeue pos, [32*nshards] fresh ids
    uint64_t* dq_word;         // deque words, [queue][worker], one 128-byte line each
    uint32_t* dq_slots;        // deque slots, [queue][worker][cap]
    int32_t* ctl;              // device control words (CtlIndex * 32)
    int32_t* host_ctl;         // host-visible control (HostCtlIndex * 32), pinned memory
    gtap_error_report* report; // host-mapped error report
    gtap_job* jobs;            // job table
    uint64_t* stats;           // [worker][ST_COUNT]
    uint64_t* prof_having;     // [worker][prof_max] ns stamps: even = has tasks, odd = ran dry
    uint64_t* prof_working;    // [worker][prof_max] ns stamps: even = round start, odd = round end
    int32_t* prof_counts;      // [worker][prof_max] tasks in the round of each working stamp
    uint64_t* prof_phase;      // [worker][PH_COUNT] shader-clock cycles per round phase (profiling build)
    uint64_t timeout_ticks;    // watchdog, wall-clock ticks (100 MHz)
    uint32_t stride;           // record stride in bytes (multiple of 16)
    uint32_t nrecs;            // task ids [0, nrecs); ids [nrecs, nrecs + nanchors) are anchors
    uint32_t nanchors;
    uint32_t ncells;           // cells per ring (power of two)
    uint32_t nshards;          // id rings
    uint32_t cap;              // deque capacity (power of two)
    uint32_t nqueues;          // provisioned deque planes
    uint32_t nworkers;         // waves in this launch
    uint32_t steal_batch;      // <= 64
    uint32_t probe_lanes;      // <= 64
    uint32_t term_mode;        // gtap_term_mode
    uint32_t njobs;            // jobs of this launch (JOBS / QUIESCE modes)
    uint32_t job_ring;         // job-table entries (anchors are job index % job_ring)
    uint32_t admit_budget;     // served: 64-id units the reservation ledger may hand out (0: no ledger)
    uint32_t priv_cap;         // ids a wave may park in its private HBM stack (<= kPrivCap)
    uint32_t* job_units;       // served: reserved 64-id units per job-table slot (device copy)
    uint32_t max_live_jobs;    // admission cap on concurrently running jobs
    uint32_t static_jobs;      // jobs [0, static_jobs) start on workers [0, static_jobs), by index
    uint32_t static_ids;       // ids [64 + w * static_ids, 64 + (w + 1) * static_ids) are worker w's
                               // at launch start (0: none)
    uint32_t prof_max;         // samples per worker; 0 = profiling off
    uint32_t idle_level;       // deepest idle backoff level (1..6: s_sleep 1 .. 32)
    uint32_t gq_cap;           // global-queue scheduler: ring slots per EPAQ queue (power of two)
    uint32_t* gq_slots;        // global-queue scheduler: [queue][gq_cap] task ids
    uint64_t* gq_ctl;          // global-queue scheduler: [queue][head, reserved, committed] * 16
    uint64_t* ovf_cells;       // deque overflow rings: [queue][ovf_cap] cells (nullptr: none)
    uint64_t* ovf_ctl;         // [queue * 32] head, [queue * 32 + 16] tickets (positions, never reset)
    uint32_t* ovf_epoch;       // [64 * 32]: per group of waves (me % 64), bumped after every push to a ring
    uint32_t ovf_cap;          // cells per overflow ring (power of two)
    uint32_t epaq_publish;     // 1: non-batch EPAQ queues publish their whole private window each round
    uint32_t xcd_steal;        // 1: even probe lanes pick victims on the thief's XCD
    uint32_t xcd_pad;
};

// Round phases of the thread-mode loop, timed in the profiling build with the shader clock
// (s_memtime): acquiring a batch (pop, admission, steal, ring), loading the batch's records
// (waited for in this build), running the task bodies, issuing the round's record writes,
// waiting for them to complete, the join atomics, appending continuations and publishing, and
// idle tries that found nothing. PH_ROUNDS counts the rounds.
enum PhaseIndex : int {
    PH_ACQUIRE = 0, PH_LOAD, PH_RUN, PH_STAGE, PH_WAIT, PH_JOIN, PH_PUBLISH, PH_IDLE, PH_ROUNDS,
    PH_COUNT = 16
};

// Timeline sample (profiling on): ns since an arbitrary epoch, from the 100 MHz wall clock.
__device__ __forceinline__ void prof_stamp(uint64_t* base, int32_t* counts, uint32_t max, uint32_t worker,
                                           uint32_t& idx, int32_t count) {
    if (idx < max) {
        const size_t at = (size_t)worker * max + idx;
        base[at] = wall_clock64() * 10ull;
        if (counts) counts[at] = count;
        idx += 1;
    }
}

// Record header (16 bytes) at the start of every task record and anchor.
//   word0: parent id (-1 = detached / none)
//   word1: byte offset of the result slot inside the parent record
//   word2: join counter (children still running)
//   word3: control word = func id [0..11] | state [12..23] | queue [24..27]
// The join counter and the control word share one 8-byte word, so the child that takes the
// counter to zero learns the parent's resume queue from the same atomic.
//
// Additive join (thread-mode lane tasks). A parent that only sums its children's integer
// results takes them through its join word instead of result slots. Its suspended header is
//   word1: result offset [0..15] | state [16..23] | func [24..30] | 1 << 31 (this layout)
//   word2..3, one 8-byte word: count [0..11] | resume queue [12..15] | sum [16..63], signed
// and each child, whose own result offset is kResSum, adds (value << 16) - 1 with the atomic
// that counts it down: one memory operation per child instead of a store and an atomic.
// Bounds: at most kSumMaxCount children per join, |value| < kSumMaxValue (so |sum| < 2^47),
// func <= kSumMaxFunc, state <= kSumMaxState; the runtime reports anything outside them.
constexpr uint32_t kHdrBytes = 16;
constexpr uint32_t kHdrJoin = 8;  // byte offset of the {join, control} 8-byte word
constexpr uint32_t kNoParent = 0xFFFFFFFFu;
constexpr uint32_t kMaxFunc = 0xFFF;
constexpr uint32_t kMaxState = 0xFFF;
constexpr uint32_t kResSum = 0xFFFFu;          // result offset of a child of an additive join
constexpr uint32_t kSumLayout = 0x80000000u;   // word1 flag of a parent waiting in an additive join
constexpr uint32_t kSumMaxCount = 0xFFFu;
constexpr uint32_t kSumMaxFunc = 0x7Fu;
constexpr uint32_t kSumMaxState = 0xFFu;
constexpr int64_t kSumMaxValue = (int64_t)1 << 35;

__device__ __forceinline__ uint32_t hdr_word0(uint32_t func, uint32_t state, uint32_t queue) {
    return (func & 0xFFFu) | ((state & 0xFFFu) << 12) | ((queue & 0xFu) << 24);
}
__device__ __forceinline__ uint32_t hdr_func(uint32_t w0) { return w0 & 0xFFFu; }
__device__ __forceinline__ uint64_t hdr_lo(uint32_t parent, uint32_t res_off) {
    return (uint64_t)parent | ((uint64_t)res_off << 32);
}
__device__ __forceinline__ uint64_t hdr_hi(uint32_t join, uint32_t ctl) { return (uint64_t)join | ((uint64_t)ctl << 32); }
__device__ __forceinline__ uint32_t hdr_state(uint32_t w0) { return (w0 >> 12) & 0xFFFu; }
__device__ __forceinline__ uint32_t hdr_queue(uint32_t w0) { return (w0 >> 24) & 0xFu; }

// ------------------------------------------------------------------ coherent access

__device__ __forceinline__ uint32_t ld32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t ld32(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st32(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st32(int32_t* p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The same agent-coherent access for a pointer known to address device (global) memory, such
// as key arrays whose addresses travel in task frames: the compiler cannot tell the address
// space of such a pointer and would use a flat instruction, which also occupies the LDS
// counter, so the LDS waits of the code around it wait for these loads too. Never pass an LDS
// or stack address.
__device__ __forceinline__ uint32_t gld32(const uint32_t* p) {
    return __hip_atomic_load((const __attribute__((address_space(1))) uint32_t*)p, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gst32(uint32_t* p, uint32_t v) {
    __hip_atomic_store((__attribute__((address_space(1))) uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t ld32_sys(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t ld32_sys(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint64_t ld64_sys(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st32_sys(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st32_sys(int32_t* p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st64_sys(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename T>
__device__ __forceinline__ T atomic_add(T* p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Fetch-add whose result is not used: the wave does not wait for it to return.
template <typename T>
__device__ __forceinline__ void atomic_add_noret(T* p, T v) {
    (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T atomic_sub(T* p, T v) {
    return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ bool atomic_cas(T* p, T& expected, T desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until every memory operation this wave issued has completed. After it, the wave's
// sc1 stores are visible device-wide, so the next atomic may publish them.
__device__ __forceinline__ void complete_stores() { __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// A program may ask for a minimum number of waves per SIMD (P::kWavesPerSimd): the compiler
// then keeps the kernel's registers within that occupancy, spilling if it must.
template <class P, class = void>
struct WavesPerSimd {
    static constexpr int value = 1;
};
template <class P>
struct WavesPerSimd<P, decltype((void)P::kWavesPerSimd)> {
    static constexpr int value = P::kWavesPerSimd;
};

// Workgroup barrier that orders LDS accesses only. __syncthreads() also waits for every global
// load and store of the calling wave to complete (a workgroup-scope fence over all address
// spaces); this one leaves them in flight, so a tile loop can keep its prefetches and its
// output stores outstanding across the barriers that hand its LDS tiles between waves.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
__device__ __forceinline__ void compiler_barrier() { __asm__ volatile("" ::: "memory"); }

// 16-byte device-coherent access to task records. A raw buffer access over the record pool
// with the sc1 cache policy is the same device-scope coherence as the 8-byte agent-scope
// relaxed atomics above (global_load/store ... sc1), at twice the width: one transaction per
// 16 bytes instead of one per 8. Offsets are 32-bit bytes from the pool base, so the pool
// (records and job anchors) stays below 4 GiB. Each call is followed by a compiler barrier so
// no global access is moved across it.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kCachePolicySc1 = 16;            // buffer instruction cache-policy operand: sc1
constexpr uint32_t kBufferRsrcWord3 = 0x00020000u;  // raw (untyped) buffer, gfx9 family
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pool_rsrc(uint8_t* base) {
    return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)0xFFFFFFFF, (int)kBufferRsrcWord3);
}
__device__ __forceinline__ u32x4 ld128(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, kCachePolicySc1);
    compiler_barrier();
    return v;
}
__device__ __forceinline__ void st128(__amdgpu_buffer_rsrc_t rs, uint32_t off, u32x4 v) {
    compiler_barrier();
    __builtin_amdgcn_raw_buffer_store_b128(v, rs, (int)off, 0, kCachePolicySc1);
    compiler_barrier();
}
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u32x2 ld64b(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off, 0, kCachePolicySc1);
    compiler_barrier();
    return v;
}
__device__ __forceinline__ void st64b(__amdgpu_buffer_rsrc_t rs, uint32_t off, u32x2 v) {
    compiler_barrier();
    __builtin_amdgcn_raw_buffer_store_b64(v, rs, (int)off, 0, kCachePolicySc1);
    compiler_barrier();
}
__device__ __forceinline__ uint32_t ld32b(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
    const uint32_t v = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, kCachePolicySc1);
    compiler_barrier();
    return v;
}
__device__ __forceinline__ void st32b(__amdgpu_buffer_rsrc_t rs, uint32_t off, uint32_t v) {
    compiler_barrier();
    __builtin_amdgcn_raw_buffer_store_b32(v, rs, (int)off, 0, kCachePolicySc1);
    compiler_barrier();
}

// The first FW frame words of the record at byte offset `rec_off`, to and from registers,
// in 16-byte pieces.
template <int FW>
__device__ __forceinline__ void frame_words_st(__amdgpu_buffer_rsrc_t rs, uint32_t rec_off, const uint32_t* f) {
#pragma unroll
    for (int i = 0; i + 4 <= FW; i += 4) st128(rs, rec_off + kHdrBytes + 4 * i, u32x4{f[i], f[i + 1], f[i + 2], f[i + 3]});
    constexpr int base = FW - FW % 4;
    if constexpr (FW % 4 >= 2) st64b(rs, rec_off + kHdrBytes + 4 * base, u32x2{f[base], f[base + 1]});
    if constexpr (FW % 2 == 1) st32b(rs, rec_off + kHdrBytes + 4 * (FW - 1), f[FW - 1]);
}
template <int FW>
__device__ __forceinline__ void frame_words_ld(__amdgpu_buffer_rsrc_t rs, uint32_t rec_off, uint32_t* f) {
#pragma unroll
    for (int i = 0; i + 4 <= FW; i += 4) {
        const u32x4 v = ld128(rs, rec_off + kHdrBytes + 4 * i);
        f[i] = v.x;
        f[i + 1] = v.y;
        f[i + 2] = v.z;
        f[i + 3] = v.w;
    }
    constexpr int base = FW - FW % 4;
    if constexpr (FW % 4 >= 2) {
        const u32x2 v = ld64b(rs, rec_off + kHdrBytes + 4 * base);
        f[base] = v.x;
        f[base + 1] = v.y;
    }
    if constexpr (FW % 2 == 1) f[FW - 1] = ld32b(rs, rec_off + kHdrBytes + 4 * (FW - 1));
}

// Typed coherent loads/stores of 4- or 8-byte aligned fields inside records.
template <typename T>
__device__ __forceinline__ T ldT(const void* p) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte fields only");
    if constexpr (sizeof(T) == 4) {
        uint32_t v = ld32(reinterpret_cast<const uint32_t*>(p));
        return __builtin_bit_cast(T, v);
    } else {
        uint64_t v = ld64(reinterpret_cast<const uint64_t*>(p));
        return __builtin_bit_cast(T, v);
    }
}
template <typename T>
__device__ __forceinline__ void stT(void* p, T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte fields only");
    if constexpr (sizeof(T) == 4) {
        st32(reinterpret_cast<uint32_t*>(p), __builtin_bit_cast(uint32_t, v));
    } else {
        st64(reinterpret_cast<uint64_t*>(p), __builtin_bit_cast(uint64_t, v));
    }
}
// Copy `nbytes` (multiple of 4, 4-byte aligned) with coherent accesses.
__device__ __forceinline__ void copy_words(void* dst, const void* src, int nbytes) {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
    for (int i = 0; i < nbytes / 4; ++i) st32(d + i, s[i]);
}

// ------------------------------------------------------------------ wave helpers

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ uint64_t active_mask() { return __ballot(1); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// number of set bits of `mask` below this lane
__device__ __forceinline__ int rank_in(uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ int popc(uint64_t m) { return __popcll(m); }
__device__ __forceinline__ int first_lane(uint64_t m) { return __ffsll((long long)m) - 1; }
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int32_t shfl(int32_t v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ uint32_t shfl(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, kWave); }
__device__ __forceinline__ uint64_t shfl(uint64_t v, int src) {
    uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, kWave);
    uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, kWave);
    return ((uint64_t)hi << 32) | lo;
}

// Volatile accessors for wave-private LDS state (the wave contexts, which live in __shared__
// arrays only). Lanes of one wave update this state inside divergent code; volatile keeps the
// compiler from reusing a value loaded before another lane's store. The pointer is cast to the
// LDS address space: through a generic pointer a volatile access becomes a flat instruction,
// which the compiler follows with a wait for every outstanding global load and store of the
// wave (s_waitcnt vmcnt(0)); an LDS access waits for nothing but itself.
template <typename T>
__device__ __forceinline__ T vld(const T* p) {
    return *(const volatile __attribute__((address_space(3))) T*)(p);
}
template <typename T>
__device__ __forceinline__ void vst(T* p, T v) {
    *(volatile __attribute__((address_space(3))) T*)(p) = v;
}

__device__ __forceinline__ uint64_t now_ticks() { return wall_clock64(); }

// Serve mode. The host's control words sit in pinned host memory; a system-scope read of
// them crosses the host bus and takes microseconds. One worker (worker 0, one lane) reads
// them, at most once per kHostPollTicks, at the top of each of its scheduler iterations,
// busy or idle, and mirrors them into device control words that the other waves read.
// A shared "last poll" word claimed by CAS would be read (and CAS'd) by every idle wave on
// every iteration: thousands of waves on one line slow every access mapped to its memory
// channel, including the running jobs' own.
// With a reservation ledger (dev.job_units), the poller also copies each newly available job's
// reservation into device memory before it publishes the count, at most kPollJobs per poll,
// so admission never reads the host's job table.
constexpr uint64_t kHostPollTicks = 100;  // 1 us of the 100 MHz wall clock
constexpr int32_t kPollJobs = 64;
__device__ __forceinline__ void host_poll(const GtapDev& dev, uint64_t& last) {
    const uint64_t now = now_ticks();
    if (now - last < kHostPollTicks) return;
    last = now;
    int32_t* ctl = dev.ctl;
    const int32_t stop = ld32_sys(dev.host_ctl + HCTL_STOP * 32);
    complete_stores();  // the stop flag is read before the job count (the host writes them in the other order)
    int32_t avail = ld32_sys(dev.host_ctl + HCTL_JOB_AVAIL * 32);
    if (dev.job_units != nullptr) {
        const int32_t seen = ld32(ctl + CTL_HOST_AVAIL * 32);
        if (avail - seen > kPollJobs) avail = seen + kPollJobs;
        for (int32_t j = seen; j < avail; ++j) {
            const uint32_t slot = (uint32_t)j % dev.job_ring;
            const uint64_t ids = ld64_sys(&dev.jobs[slot].reserve_ids);
            st32(dev.job_units + slot, (uint32_t)((ids + 63u) / 64u));
        }
        complete_stores();  // the reservations are in place before the jobs become visible
    }
    // one writer and a non-decreasing host count: a plain store keeps the mirror monotonic
    st32(ctl + CTL_HOST_AVAIL * 32, avail);
    if (stop && avail == ld32_sys(dev.host_ctl + HCTL_JOB_AVAIL * 32)) {
        complete_stores();  // a wave that sees the stop flag sees the final job count
        st32(ctl + CTL_HOST_STOP * 32, 1);
    }
}

__device__ __forceinline__ uint32_t xorshift32(uint32_t& s) {
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}

// ------------------------------------------------------------------ errors

// First error wins; the report goes to host-mapped memory; every wave then leaves its
// scheduler loop (the kernel ends normally instead of trapping).
// Takes scalars only, so the kernel argument block is never copied to scratch.
__device__ __noinline__ void record_error(int32_t* ctl, gtap_error_report* report, int code, int tid, int queue,
                                          int value, int limit, int line);

#define GTAP_DEVICE_ERROR(dev, code, tid, queue, value, limit)                                             \
    ::gtap::record_error((dev).ctl, (dev).report, (code), (int)(tid), (int)(queue), (int)(value), (int)(limit), \
                         __LINE__)

__device__ __noinline__ void record_error(int32_t* ctl, gtap_error_report* report, int code, int tid, int queue,
                                          int value, int limit, int line) {
    int32_t expected = 0;
    if (atomic_cas(ctl + CTL_ERROR * 32, expected, (int32_t)code)) {
        gtap_error_report* r = report;
        if (r != nullptr) {
            st32_sys(&r->code, code);
            st32_sys(&r->src_line, line);
            st32_sys(&r->block_idx, (int32_t)blockIdx.x);
            st32_sys(&r->thread_idx, (int32_t)threadIdx.x);
            st32_sys(&r->tid, tid);
            st32_sys(&r->queue_idx, queue);
            st32_sys(&r->value, value);
            st32_sys(&r->limit, limit);
            complete_stores();
            st32_sys(&r->valid, 1);
        }
    }
    st32(ctl + CTL_ABORT * 32, 1);
    // idle waves watch only the done word
    __hip_atomic_fetch_or(ctl + CTL_DONE * 32, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace gtap
