// fib in the #pragma gtap dialect: each call spawns two tasks and waits for both.
#include <cuda_runtime.h>
#include "gtap_thread.cuh"

__device__ long long d_fib_result;

#pragma gtap function
__device__ long long fib(int n) {
    if (n < 2) return n;
    long long a, b;
    #pragma gtap task
    a = fib(n - 1);
    #pragma gtap task
    b = fib(n - 2);
    #pragma gtap taskwait
    return a + b;
}

__global__ void fib_kernel(int n) {
    #pragma gtap entry
    d_fib_result = fib(n);
}
