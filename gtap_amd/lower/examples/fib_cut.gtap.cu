// The same program as fib_epaq.gtap.cu on one queue: the comparison point of the EPAQ form.
#include <cuda_runtime.h>
#include "gtap_thread.cuh"

__device__ long long d_result;
__device__ __constant__ int d_cutoff;

// below the cutoff: plain recursion, as the reference's EPAQ benchmark does
// (evaluation/3-epaq/fib/fib_queue_3.cu:9-14), so leaf tasks run long and unlike the others
__device__ long long fib_serial(int n) {
    if (n < 2) return n;
    return fib_serial(n - 1) + fib_serial(n - 2);
}

#pragma gtap function
__device__ long long fib(int n) {
    if (n < d_cutoff) return fib_serial(n);
    long long x, y;
    #pragma gtap task
    x = fib(n - 1);
    #pragma gtap task
    y = fib(n - 2);
    #pragma gtap taskwait
    return x + y;
}

__global__ void fib_kernel(int n) {
    #pragma gtap entry
    d_result = fib(n);
}
