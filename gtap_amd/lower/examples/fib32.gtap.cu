// fib with a 32-bit result, the shape of the reference's examples/fib/fib.cu: its children's
// results are only summed after the taskwait, so the lowering joins them additively.
#include <cuda_runtime.h>
#include "gtap_thread.cuh"

__device__ int d_fib32_result;

#pragma gtap function
__device__ int fib(int n) {
    if (n < 2) return n;
    int a, b;
    #pragma gtap task
    a = fib(n - 1);
    #pragma gtap task
    b = fib(n - 2);
    #pragma gtap taskwait
    return a + b;
}

__global__ void fib32_kernel(int n) {
    #pragma gtap entry
    d_fib32_result = fib(n);
}
