// n-queens in the #pragma gtap dialect: one task per legal placement in the next row,
// the counts of the children summed after a taskwait.
#include <cuda_runtime.h>
#include "gtap_thread.cuh"

__device__ long long d_nq_result;

#pragma gtap function
__device__ long long nqueens(int n, int row, unsigned cols, unsigned diag, unsigned anti) {
    if (row == n) return 1;
    unsigned avail = ((1u << n) - 1u) & ~(cols | diag | anti);
    long long counts[16];
    int k = 0;
    while (avail) {
        unsigned p = avail & (0u - avail);
        avail -= p;
        #pragma gtap task
        counts[k++] = nqueens(n, row + 1, cols | p, (diag | p) << 1, (anti | p) >> 1);
    }
    #pragma gtap taskwait
    long long s = 0;
    for (int i = 0; i < k; ++i) s += counts[i];
    return s;
}

__global__ void nqueens_kernel(int n) {
    #pragma gtap entry
    d_nq_result = nqueens(n, 0, 0u, 0u, 0u);
}
