// gtap_lowered.hpp -- what code written by gtap_amd.lower includes besides the device ABI:
//
//  * coherent access to memory that tasks share (gtap::cld / gtap::cst / gtap::coh). The
//    lowering wraps every access through a pointer inside task functions, their device
//    helpers and entry kernels in these, because this runtime publishes work without cache
//    fences: a plain store can sit in one XCD's L2 while a task on another XCD reads the
//    line from memory. They are agent-scope relaxed atomics (sc1 loads and stores), the
//    same coherence the runtime uses for its own records;
//  * the reference's low-level access helpers with its names and signatures
//    (runtime/detail/common/gtap_runtime_common.cuh:89-185: load_L2, store_L2, load_L2_acquire,
//    store_L2_release, load_L2_ptr, store_L2_ptr, store_struct_L2, atomicCAS_acquire, lock,
//    unlock), written as agent-scope atomics instead of PTX cache operators;
//  * the entry of a host-launched `__global__` kernel (the reference's `#pragma gtap entry`,
//    examples/fib/fib.cu:21-46): every thread of the launch calls entry_thread (thread mode)
//    or entry_block (block mode). Worker 0 writes the root job, every wave or workgroup runs
//    the persistent scheduler, and worker 0's lane 0 learns the root's result when the
//    launch is complete. The runtime state comes from the descriptor that gtap_initialize /
//    gtap_reset push to each lowered module (gtap_register_entry_module); one launch may
//    follow each of those calls (the reset arms the state, the launch disarms it, and a
//    launch on a disarmed state reports GTAP_ERROR_NOT_INITIALIZED and does nothing).
#pragma once

#include <type_traits>

#include "gtap_abi.hpp"
#include "gtap_block.hpp"

#ifndef GTAP_WARP_SIZE
#define GTAP_WARP_SIZE 64  // wavefront width (the reference's warp is 32 lanes)
#endif
#ifndef GTAP_BLOCK_SIZE
#define GTAP_BLOCK_SIZE 256  // threads per workgroup of an entry launch (4 wavefront workers)
#endif
#ifndef GTAP_NUM_WARPS
#define GTAP_NUM_WARPS (GTAP_BLOCK_SIZE / GTAP_WARP_SIZE)
#endif

namespace gtap {

// ------------------------------------------------------------------ coherent access

template <class T>
__device__ __forceinline__ std::remove_cv_t<T> cld(T* p) {
    using U = std::remove_cv_t<T>;
    static_assert(std::is_trivially_copyable<U>::value, "coherent access to a trivially copyable type only");
    if constexpr (sizeof(U) == 8 && alignof(U) >= 8) {
        return __builtin_bit_cast(U, __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT));
    } else if constexpr (sizeof(U) == 4 && alignof(U) >= 4) {
        return __builtin_bit_cast(U, __hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT));
    } else if constexpr (sizeof(U) == 2 && alignof(U) >= 2) {
        return __builtin_bit_cast(U, __hip_atomic_load(reinterpret_cast<const uint16_t*>(p), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT));
    } else if constexpr (sizeof(U) == 1) {
        return __builtin_bit_cast(U, __hip_atomic_load(reinterpret_cast<const uint8_t*>(p), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT));
    } else {
        static_assert(sizeof(U) % 4 == 0 && alignof(U) >= 4, "coherent access: 1, 2, 4, 8 bytes or whole 4-byte words");
        struct W {
            uint32_t w[sizeof(U) / 4];
        } v;
        const uint32_t* s = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
        for (unsigned i = 0; i < sizeof(U) / 4; ++i)
            v.w[i] = __hip_atomic_load(s + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return __builtin_bit_cast(U, v);
    }
}

template <class T>
__device__ __forceinline__ void cst(T* p, const std::remove_cv_t<T>& v) {
    using U = std::remove_cv_t<T>;
    static_assert(!std::is_const<T>::value, "store through a pointer to const");
    static_assert(std::is_trivially_copyable<U>::value, "coherent access to a trivially copyable type only");
    if constexpr (sizeof(U) == 8 && alignof(U) >= 8) {
        __hip_atomic_store(reinterpret_cast<uint64_t*>(p), __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else if constexpr (sizeof(U) == 4 && alignof(U) >= 4) {
        __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __builtin_bit_cast(uint32_t, v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else if constexpr (sizeof(U) == 2 && alignof(U) >= 2) {
        __hip_atomic_store(reinterpret_cast<uint16_t*>(p), __builtin_bit_cast(uint16_t, v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else if constexpr (sizeof(U) == 1) {
        __hip_atomic_store(reinterpret_cast<uint8_t*>(p), __builtin_bit_cast(uint8_t, v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else {
        static_assert(sizeof(U) % 4 == 0 && alignof(U) >= 4, "coherent access: 1, 2, 4, 8 bytes or whole 4-byte words");
        struct W {
            uint32_t w[sizeof(U) / 4];
        };
        const W s = __builtin_bit_cast(W, v);
        uint32_t* d = reinterpret_cast<uint32_t*>(p);
#pragma unroll
        for (unsigned i = 0; i < sizeof(U) / 4; ++i)
            __hip_atomic_store(d + i, s.w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// An lvalue in shared memory that is assigned, compound-assigned or incremented: every read
// is a coherent load and every write a coherent store (a read-modify-write is a load and a
// store, as in the plain code; it is not atomic).
template <class T>
struct Coh {
    using U = std::remove_cv_t<T>;
    T* p;
    __device__ __forceinline__ operator U() const { return cld(p); }
    __device__ __forceinline__ Coh& operator=(const U& v) {
        cst(p, v);
        return *this;
    }
    __device__ __forceinline__ Coh& operator=(const Coh& o) {
        cst(p, cld(o.p));
        return *this;
    }
#define GTAP_COH_COMPOUND(op)                          \
    template <class V>                                 \
    __device__ __forceinline__ Coh& operator op##=(const V& v) { \
        cst(p, (U)(cld(p) op v));                      \
        return *this;                                  \
    }
    GTAP_COH_COMPOUND(+)
    GTAP_COH_COMPOUND(-)
    GTAP_COH_COMPOUND(*)
    GTAP_COH_COMPOUND(/)
    GTAP_COH_COMPOUND(%)
    GTAP_COH_COMPOUND(&)
    GTAP_COH_COMPOUND(|)
    GTAP_COH_COMPOUND(^)
    GTAP_COH_COMPOUND(<<)
    GTAP_COH_COMPOUND(>>)
#undef GTAP_COH_COMPOUND
    __device__ __forceinline__ Coh& operator++() {
        U v = cld(p);
        ++v;
        cst(p, v);
        return *this;
    }
    __device__ __forceinline__ Coh& operator--() {
        U v = cld(p);
        --v;
        cst(p, v);
        return *this;
    }
    __device__ __forceinline__ U operator++(int) {
        const U v = cld(p);
        U n = v;
        ++n;
        cst(p, n);
        return v;
    }
    __device__ __forceinline__ U operator--(int) {
        const U v = cld(p);
        U n = v;
        --n;
        cst(p, n);
        return v;
    }
};
template <class T>
__device__ __forceinline__ Coh<T> coh(T& r) {
    return Coh<T>{&r};
}

// A root's return value from the job table's 8-byte result (the low sizeof(T) bytes).
template <class T>
__device__ __forceinline__ T result_as(int64_t bits) {
    static_assert(sizeof(T) <= 8 && std::is_trivially_copyable<T>::value, "task results are at most 8 bytes");
    if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, bits);
    } else if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, (uint32_t)(uint64_t)bits);
    } else if constexpr (sizeof(T) == 2) {
        return __builtin_bit_cast(T, (uint16_t)(uint64_t)bits);
    } else {
        return __builtin_bit_cast(T, (uint8_t)(uint64_t)bits);
    }
}

// ------------------------------------------------------------------ host-launched entry

// Worker 0 writes the root job into slot 0 of the job table (device memory): the job that
// the scheduler's first worker starts by index.
template <class A>
__device__ __forceinline__ void entry_write_job(const GtapDev& dev, uint32_t func, uint32_t queue, const A& root) {
    static_assert(sizeof(A) % 8 == 0 && sizeof(A) <= sizeof(((gtap_job*)nullptr)->args), "root arguments");
    gtap_job* j = dev.jobs;
    st32(&j->func, func);
    st32(&j->queue, queue);
    st64(reinterpret_cast<uint64_t*>(&j->result), 0ull);
    const uint64_t* s = reinterpret_cast<const uint64_t*>(&root);
    uint64_t* d = reinterpret_cast<uint64_t*>(j->args);
#pragma unroll
    for (unsigned i = 0; i < sizeof(A) / 8; ++i) st64(d + i, s[i]);
    complete_stores();
    st32(&j->status, (uint32_t)GTAP_JOB_READY);
    complete_stores();
}

// Worker 0 after its scheduler loop: the root's result if the job completed; disarms the
// state (a launch without gtap_reset in between then does nothing but report the error).
__device__ __forceinline__ bool entry_collect(const GtapDev& dev, int64_t& result) {
    const uint32_t status = ld32(&dev.jobs->status);
    const int64_t r = (int64_t)ld64(reinterpret_cast<const uint64_t*>(&dev.jobs->result));
    st32(dev.ctl + CTL_ARMED * 32, 0);
    if (status != (uint32_t)GTAP_JOB_DONE) return false;
    result = r;
    return true;
}

// Thread mode: each wavefront of the launch is one worker. Returns true on the one thread
// (block 0, thread 0) that holds the root's result, after the launch's work is complete.
template <class P, class A>
__device__ __forceinline__ bool entry_thread(const GtapDev& base, uint32_t func, uint32_t queue, const A& root,
                                             int64_t& result) {
    __shared__ WaveCtxOf<P> ctxs[4];
    __shared__ typename WaveSharedOf<P>::type wsh[4];
    const int lane = lane_id();
    const bool first = blockIdx.x == 0 && threadIdx.x == 0;
    if (base.ctl == nullptr || base.jobs == nullptr) return false;  // no descriptor pushed yet (no gtap_initialize)
    if (blockDim.x % kWave != 0 || blockDim.x > 256 || blockDim.y != 1 || blockDim.z != 1 || gridDim.y != 1 ||
        gridDim.z != 1) {
        if (first) GTAP_DEVICE_ERROR(base, GTAP_ERROR_INVALID_ARGUMENT, -1, -1, (int)blockDim.x, 256);
        return false;
    }
    if ((uint32_t)P::kQueues > base.nqueues || P::kFrameBytes + kHdrBytes > base.stride) {
        if (first) GTAP_DEVICE_ERROR(base, GTAP_ERROR_INVALID_ARGUMENT, -1, -1, (int)P::kFrameBytes, (int)base.stride);
        return false;
    }
    GtapDev dev = base;
    const uint32_t waves = blockDim.x / kWave;
    const uint32_t launched = gridDim.x * waves;
    if (launched < dev.nworkers) dev.nworkers = launched;  // the descriptor holds the provisioned maximum
    const uint32_t me = blockIdx.x * waves + threadIdx.x / kWave;
    if (me >= dev.nworkers) return false;
    int32_t armed = 0;
    if (lane == 0) armed = ld32(dev.ctl + CTL_ARMED * 32);
    armed = shfl(armed, 0);
    if (!armed) {
        if (first) GTAP_DEVICE_ERROR(dev, GTAP_ERROR_NOT_INITIALIZED, -1, -1, 0, 0);
        return false;
    }
    if (me == 0 && lane == 0) entry_write_job(dev, func, queue, root);
    const typename P::Args args{};
    thread_scheduler<P, kSchedWS, false>(dev, args, &ctxs[threadIdx.x / kWave], &wsh[threadIdx.x / kWave]);
    if (me == 0 && lane == 0) return entry_collect(dev, result);
    return false;
}

// Block mode: each workgroup of 256 threads is one worker.
template <class P, class A>
__device__ __forceinline__ bool entry_block(const GtapDev& base, uint32_t func, const A& root, int64_t& result) {
    __shared__ BlockCtx ctx;
    __shared__ typename P::Shared sh;
    __shared__ int32_t s_armed;
    const uint32_t t = threadIdx.x;
    const bool first = blockIdx.x == 0 && t == 0;
    if (base.ctl == nullptr || base.jobs == nullptr) return false;  // no descriptor pushed yet (no gtap_initialize)
    if (blockDim.x != kBlockThreads || blockDim.y != 1 || blockDim.z != 1 || gridDim.y != 1 || gridDim.z != 1) {
        if (first) GTAP_DEVICE_ERROR(base, GTAP_ERROR_INVALID_ARGUMENT, -1, -1, (int)blockDim.x, (int)kBlockThreads);
        return false;
    }
    if (P::kFrameBytes + kHdrBytes > base.stride) {
        if (first) GTAP_DEVICE_ERROR(base, GTAP_ERROR_INVALID_ARGUMENT, -1, -1, (int)P::kFrameBytes, (int)base.stride);
        return false;
    }
    GtapDev dev = base;
    if (gridDim.x < dev.nworkers) dev.nworkers = gridDim.x;
    const uint32_t me = blockIdx.x;
    if (me >= dev.nworkers) return false;
    if (t == 0) {
        s_armed = ld32(dev.ctl + CTL_ARMED * 32);
        if (!s_armed && first) GTAP_DEVICE_ERROR(dev, GTAP_ERROR_NOT_INITIALIZED, -1, -1, 0, 0);
        if (s_armed && me == 0) entry_write_job(dev, func, 0u, root);
    }
    __syncthreads();
    if (!s_armed) return false;
    const typename P::Args args{};
    block_scheduler<P>(dev, args, &ctx, &sh);
    if (me == 0 && t == 0) return entry_collect(dev, result);
    return false;
}

}  // namespace gtap

// ------------------------------------------------------------------ the reference's helpers
// Same names and signatures as runtime/detail/common/gtap_runtime_common.cuh:89-185. The
// reference reads and writes "at L2" with PTX cache operators, relying on one coherent L2;
// here every one is an agent-scope atomic access, coherent across the XCDs' L2 caches.

__device__ __forceinline__ int load_L2(int* ptr) { return gtap::cld(ptr); }
__device__ __forceinline__ unsigned int load_L2(unsigned int* ptr) { return gtap::cld(ptr); }
__device__ __forceinline__ uint16_t load_L2_u16t(uint16_t* ptr) { return gtap::cld(ptr); }
__device__ __forceinline__ int load_L2_acquire(int* ptr) {
    return __hip_atomic_load(ptr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void* load_L2_ptr(void** ptr) { return gtap::cld(ptr); }
template <typename T>
__device__ __forceinline__ T* load_L2_ptr(T** ptr) {
    return gtap::cld(ptr);
}
__device__ __forceinline__ void store_L2(int* ptr, int val) { gtap::cst(ptr, val); }
__device__ __forceinline__ void store_L2(unsigned int* ptr, unsigned int val) { gtap::cst(ptr, val); }
__device__ __forceinline__ void store_L2_u16t(uint16_t* ptr, uint16_t val) { gtap::cst(ptr, val); }
__device__ __forceinline__ void store_L2_release(int* ptr, int val) {
    __hip_atomic_store(ptr, val, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_L2_ptr(void** ptr, void* val) { gtap::cst(ptr, val); }
template <typename T>
__device__ __forceinline__ void store_L2_ptr(T** ptr, T* val) {
    gtap::cst(ptr, val);
}
template <typename T>
__device__ __forceinline__ void store_struct_L2(T* dst, const T& src) {
    gtap::cst(dst, src);
}
__device__ __forceinline__ int atomicCAS_acquire(int* addr, int expected, int desired) {
    __hip_atomic_compare_exchange_strong(addr, &expected, desired, __ATOMIC_ACQUIRE, __ATOMIC_ACQUIRE,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expected;
}
__device__ __forceinline__ void lock(int* lock_var) {
    while (atomicCAS_acquire(lock_var, 0, 1) != 0) __builtin_amdgcn_s_sleep(1);
}
__device__ __forceinline__ void unlock(int* lock_var) {
    __hip_atomic_store(lock_var, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
// Lane and worker index inside the workgroup (the reference's warp is 32 lanes, a wavefront 64).
__device__ __forceinline__ unsigned int get_lane_id() { return threadIdx.x & (GTAP_WARP_SIZE - 1); }
__device__ __forceinline__ unsigned int get_warp_id_in_block() { return threadIdx.x / GTAP_WARP_SIZE; }
