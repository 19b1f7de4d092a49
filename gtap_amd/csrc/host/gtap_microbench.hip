// gtap_microbench.hip -- measured peaks of this GPU, for the roofline fractions the
// benchmark reports: HBM streaming-copy bandwidth and agent-scope atomic throughput
// (relaxed fetch-add, each lane on its own 64-byte-aligned word of a 256 MiB region, and a
// contended form over a few words);
// random-access rates (plain loads and fetch-min over a table, the BFS request ceiling), and
// the cost of the publication orders the scheduler could use (gtap_bench_fences).
#include <hip/hip_runtime.h>

#include "gtap_launch.hpp"

namespace {

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// Streaming copy with U 16-byte loads in flight per thread before their stores; NT: the
// loads and stores are non-temporal (no reuse expected, so they need not be kept in cache).
template <int U, bool NT>
__global__ void __launch_bounds__(256) copy_kernel(const v4u* __restrict__ a, v4u* __restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (; i + (size_t)(U - 1) * stride < n; i += (size_t)U * stride) {
        v4u v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = NT ? __builtin_nontemporal_load(a + i + u * stride) : a[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (NT) __builtin_nontemporal_store(v[u], b + i + u * stride);
            else b[i + u * stride] = v[u];
        }
    }
    for (; i < n; i += stride) b[i] = a[i];
}

__global__ void atomic_kernel(uint32_t* buf, int iters, uint32_t nslots) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = 0; i < iters; ++i) {
        const uint32_t slot = (g * 16u + (uint32_t)i * 977u) % nslots;
        __hip_atomic_fetch_add(buf + (size_t)slot * 16, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Random 4-byte accesses to a table of nwords words, 8 independent ones in flight per thread
// per step (as the BFS expansion issues its depth probes): ATOMIC 0 plain loads, 1 agent-scope
// relaxed fetch-min (its result used). The index stream is a per-thread hash, so nothing but
// the table is read; each thread writes one word so the loads are kept.
template <int ATOMIC>
__global__ void __launch_bounds__(256) random_kernel(uint32_t* table, uint32_t nwords, int steps, uint32_t* sink) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t x = g * 0x9E3779B9u + 0x7F4A7C15u;
    uint32_t acc = 0;
    for (int s = 0; s < steps; ++s) {
        uint32_t idx[8], v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            x ^= x << 13;
            x ^= x >> 17;
            x ^= x << 5;
            idx[j] = x % nwords;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if constexpr (ATOMIC) v[j] = __hip_atomic_fetch_min(table + idx[j], g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else v[j] = table[idx[j]];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) acc += v[j];
    }
    sink[g] = acc;
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

// Streaming copy of `bytes` (read + write counted): the best of `reps` launches of each copy
// variant (1 or 4 loads in flight per thread, plain or non-temporal accesses, two grid sizes).
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
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const v4u* src = (const v4u*)a;
    v4u* dst = (v4u*)b;
    const size_t n = bytes / 16;
    double best = 0;
    for (int variant = 0; variant < 8; ++variant) {
        const dim3 grid((unsigned)cus * ((variant & 4) ? 8u : 32u));
        for (int r = 0; r < reps + 1; ++r) {
            GTAP_HIP_TRY(hipEventRecord(e0, nullptr));
            switch (variant & 3) {
                case 0: hipLaunchKernelGGL((copy_kernel<1, false>), grid, dim3(256), 0, nullptr, src, dst, n); break;
                case 1: hipLaunchKernelGGL((copy_kernel<1, true>), grid, dim3(256), 0, nullptr, src, dst, n); break;
                case 2: hipLaunchKernelGGL((copy_kernel<4, false>), grid, dim3(256), 0, nullptr, src, dst, n); break;
                default: hipLaunchKernelGGL((copy_kernel<4, true>), grid, dim3(256), 0, nullptr, src, dst, n); break;
            }
            GTAP_HIP_TRY(hipEventRecord(e1, nullptr));
            GTAP_HIP_TRY(hipEventSynchronize(e1));
            float ms = 0;
            GTAP_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            const double g = 2.0 * (double)bytes / (ms * 1e6);
            if (r > 0 && g > best) best = g;  // the first launch of a variant warms up
        }
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

// Contended agent-scope relaxed atomic adds: every lane adds to one of `words` words (64 bytes
// apart), lanes spread over them by global index; best of `reps`.
extern "C" int gtap_bench_atomics_contended(int reps, uint32_t words, double* gops) {
    if (reps < 1 || !gops || words < 1 || words > (1u << 22)) return gtap::host::invalid_argument("gtap_bench_atomics_contended");
    uint32_t* buf = nullptr;
    GTAP_HIP_TRY(hipMalloc(&buf, (size_t)words * 64));
    GTAP_HIP_TRY(hipMemset(buf, 0, (size_t)words * 64));
    hipEvent_t e0, e1;
    GTAP_HIP_TRY(hipEventCreate(&e0));
    GTAP_HIP_TRY(hipEventCreate(&e1));
    const int blocks = 256 * 8, threads = 256, iters = 16;
    double best = 0;
    for (int r = 0; r < reps + 1; ++r) {
        GTAP_HIP_TRY(hipEventRecord(e0, nullptr));
        hipLaunchKernelGGL(atomic_kernel, dim3(blocks), dim3(threads), 0, nullptr, buf, iters, words);
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

// Random-access rates over a table of table_bytes (the BFS depth array is 16 MiB at R-MAT
// scale 22): out[0] random 4-byte loads, out[1] random agent-scope fetch-min, in G accesses/s;
// best of `reps` each. The request-rate ceiling of the BFS expansion, beside the byte one.
extern "C" int gtap_bench_random_access(uint64_t table_bytes, int reps, double* out) {
    if (reps < 1 || !out || table_bytes < 4096 || table_bytes > (4ull << 30))
        return gtap::host::invalid_argument("gtap_bench_random_access: table of 4 KiB .. 4 GiB");
    const uint32_t nwords = (uint32_t)(table_bytes / 4);
    const int blocks = 256 * 16, threads = 256, steps = 16;
    uint32_t *table = nullptr, *sink = nullptr;
    GTAP_HIP_TRY(hipMalloc(&table, (size_t)nwords * 4));
    GTAP_HIP_TRY(hipMalloc(&sink, (size_t)blocks * threads * 4));
    GTAP_HIP_TRY(hipMemset(table, 0x3f, (size_t)nwords * 4));
    hipEvent_t e0, e1;
    GTAP_HIP_TRY(hipEventCreate(&e0));
    GTAP_HIP_TRY(hipEventCreate(&e1));
    for (int mode = 0; mode < 2; ++mode) {
        double best = 0;
        for (int r = 0; r < reps + 1; ++r) {
            GTAP_HIP_TRY(hipEventRecord(e0, nullptr));
            if (mode == 0) hipLaunchKernelGGL(random_kernel<0>, dim3(blocks), dim3(threads), 0, nullptr, table, nwords, steps, sink);
            else hipLaunchKernelGGL(random_kernel<1>, dim3(blocks), dim3(threads), 0, nullptr, table, nwords, steps, sink);
            GTAP_HIP_TRY(hipEventRecord(e1, nullptr));
            GTAP_HIP_TRY(hipEventSynchronize(e1));
            float ms = 0;
            GTAP_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            const double g = (double)blocks * threads * steps * 8 / (ms * 1e6);
            if (r > 0 && g > best) best = g;
        }
        out[mode] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(table);
    (void)hipFree(sink);
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
