// gtap_microbench.hip -- measured peaks of this GPU, for the roofline fractions the
// benchmark reports: HBM streaming-copy bandwidth and agent-scope atomic throughput
// (relaxed fetch-add, each lane on its own 64-byte-aligned word of a 256 MiB region);
// and the cost of the publication orders the scheduler could use (gtap_bench_fences).
#include <hip/hip_runtime.h>

#include "gtap_launch.hpp"

namespace {

__global__ void copy_kernel(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

__global__ void atomic_kernel(uint32_t* buf, int iters, uint32_t nslots) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = 0; i < iters; ++i) {
        const uint32_t slot = (g * 16u + (uint32_t)i * 977u) % nslots;
        __hip_atomic_fetch_add(buf + (size_t)slot * 16, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One publication per iteration per wave: a data store, then an order, then a relaxed
// agent-scope flag store. MODE 0: agent-scope release fence; 1: relaxed flag load followed by an
// agent-scope acquire fence; 2: wait for the wave's own stores to complete (s_waitcnt vmcnt(0)),
// the order the scheduler uses.
template <int MODE>
__global__ void __launch_bounds__(256) fence_kernel(uint32_t* buf, int iters) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if ((threadIdx.x & 63u) != 0u) return;
    uint32_t* p = buf + (size_t)wave * 16u;
    uint32_t acc = 0;
    for (int i = 0; i < iters; ++i) {
        __hip_atomic_store(p + 1, (uint32_t)i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if constexpr (MODE == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if constexpr (MODE == 1) {
            acc += __hip_atomic_load(p + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        if constexpr (MODE == 2) __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __hip_atomic_store(p, (uint32_t)i + acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace

using gtap::host::hip_status;

// Streaming copy of `bytes` (read + write counted); best of `reps`.
extern "C" int gtap_bench_hbm_copy(uint64_t bytes, int reps, double* gbps) {
    if (bytes < (1u << 20) || reps < 1 || !gbps) return gtap::host::invalid_argument("gtap_bench_hbm_copy");
    bytes &= ~(uint64_t)15;
    void *a = nullptr, *b = nullptr;
    GTAP_HIP_TRY(hipMalloc(&a, bytes));
    GTAP_HIP_TRY(hipMalloc(&b, bytes));
    GTAP_HIP_TRY(hipMemset(a, 1, bytes));
    hipEvent_t e0, e1;
    GTAP_HIP_TRY(hipEventCreate(&e0));
    GTAP_HIP_TRY(hipEventCreate(&e1));
    double best = 0;
    for (int r = 0; r < reps + 1; ++r) {
        GTAP_HIP_TRY(hipEventRecord(e0, nullptr));
        hipLaunchKernelGGL(copy_kernel, dim3(256 * 32), dim3(256), 0, nullptr, (const uint4*)a, (uint4*)b, bytes / 16);
        GTAP_HIP_TRY(hipEventRecord(e1, nullptr));
        GTAP_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        GTAP_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        const double g = 2.0 * (double)bytes / (ms * 1e6);
        if (r > 0 && g > best) best = g;  // first launch warms up
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    *gbps = best;
    return 0;
}

// Agent-scope relaxed atomic adds on distinct 64-byte-aligned words; best of `reps`.
extern "C" int gtap_bench_atomics(int reps, double* gops) {
    if (reps < 1 || !gops) return gtap::host::invalid_argument("gtap_bench_atomics");
    const uint32_t nslots = (256u << 20) / 64u;
    uint32_t* buf = nullptr;
    GTAP_HIP_TRY(hipMalloc(&buf, (size_t)nslots * 64));
    GTAP_HIP_TRY(hipMemset(buf, 0, (size_t)nslots * 64));
    hipEvent_t e0, e1;
    GTAP_HIP_TRY(hipEventCreate(&e0));
    GTAP_HIP_TRY(hipEventCreate(&e1));
    const int blocks = 256 * 32, threads = 256, iters = 64;
    double best = 0;
    for (int r = 0; r < reps + 1; ++r) {
        GTAP_HIP_TRY(hipEventRecord(e0, nullptr));
        hipLaunchKernelGGL(atomic_kernel, dim3(blocks), dim3(threads), 0, nullptr, buf, iters, nslots);
        GTAP_HIP_TRY(hipEventRecord(e1, nullptr));
        GTAP_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        GTAP_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        const double g = (double)blocks * threads * iters / (ms * 1e6);
        if (r > 0 && g > best) best = g;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    *gops = best;
    return 0;
}

// Time per publication per wave, in ns, with `waves` waves each doing `iters` publications
// concurrently: out_ns[0] release fence, out_ns[1] acquire fence, out_ns[2] store completion wait.
extern "C" int gtap_bench_fences(uint32_t waves, int iters, double* out_ns) {
    if (waves < 4 || waves > (1u << 20) || (waves & 3u) || iters < 1 || !out_ns)
        return gtap::host::invalid_argument("gtap_bench_fences: waves must be a multiple of 4 in [4, 2^20]");
    uint32_t* buf = nullptr;
    GTAP_HIP_TRY(hipMalloc(&buf, (size_t)waves * 64));
    GTAP_HIP_TRY(hipMemset(buf, 0, (size_t)waves * 64));
    hipEvent_t e0, e1;
    GTAP_HIP_TRY(hipEventCreate(&e0));
    GTAP_HIP_TRY(hipEventCreate(&e1));
    const dim3 grid(waves / 4), block(256);
    for (int mode = 0; mode < 3; ++mode) {
        double best = 1e300;
        for (int r = 0; r < 3; ++r) {
            GTAP_HIP_TRY(hipEventRecord(e0, nullptr));
            if (mode == 0) hipLaunchKernelGGL(fence_kernel<0>, grid, block, 0, nullptr, buf, iters);
            if (mode == 1) hipLaunchKernelGGL(fence_kernel<1>, grid, block, 0, nullptr, buf, iters);
            if (mode == 2) hipLaunchKernelGGL(fence_kernel<2>, grid, block, 0, nullptr, buf, iters);
            GTAP_HIP_TRY(hipEventRecord(e1, nullptr));
            GTAP_HIP_TRY(hipEventSynchronize(e1));
            float ms = 0;
            GTAP_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0 && ms < best) best = ms;
        }
        out_ns[mode] = best * 1e6 / iters;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    return 0;
}
