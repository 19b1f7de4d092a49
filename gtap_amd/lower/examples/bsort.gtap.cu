// Block-mode merge sort in the #pragma gtap dialect: every task runs on a whole 256-thread
// workgroup. A leaf of at most 256 keys is sorted in workgroup memory by an odd-even
// transposition network; a larger range spawns its two halves from thread 0, waits, and
// merges them with every thread placing keys by their rank in the other half.
#include <cuda_runtime.h>
#include "gtap_block.cuh"

__device__ unsigned* g_keys;
__device__ unsigned* g_tmp;

// rank of v among x[0, n): keys below it (strict) or not above it (ties go to the left run)
__device__ int rank_of(unsigned* x, int n, unsigned v, bool strict) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        unsigned m = x[mid];
        if (strict ? (m < v) : (m <= v)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

#pragma gtap function
__device__ void bsort(unsigned* a, unsigned* t, int n) {
    __shared__ unsigned tile[256];
    if (n <= 256) {
        int i = threadIdx.x;
        tile[i] = i < n ? a[i] : 0xFFFFFFFFu;
        __syncthreads();
        for (int round = 0; round < 256; ++round) {
            int j = 2 * i + (round & 1);
            if (i < 128 && j + 1 < 256 && tile[j] > tile[j + 1]) {
                unsigned x = tile[j];
                tile[j] = tile[j + 1];
                tile[j + 1] = x;
            }
            __syncthreads();
        }
        if (i < n) a[i] = tile[i];
        return;
    }
    int h = n / 2;
    if (threadIdx.x == 0) {
        #pragma gtap task
        bsort(a, t, h);
        #pragma gtap task
        bsort(a + h, t + h, n - h);
    }
    #pragma gtap taskwait
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        unsigned v = a[k];
        int pos = k < h ? k + rank_of(a + h, n - h, v, true) : (k - h) + rank_of(a, h, v, false);
        t[pos] = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) a[k] = t[k];
}

__global__ void bsort_kernel(int n) {
    #pragma gtap entry
    bsort(g_keys, g_tmp, n);
}
