// fib with EPAQ queues in the #pragma gtap dialect: below a cutoff that the host sets in a
// __constant__ variable the call is sequential; a child that will run sequentially goes to
// queue 1, one that will spawn again to queue 0, and a parent resumes on queue 2, so each
// wavefront batch holds tasks of one execution path.
#include <cuda_runtime.h>
#include "gtap_thread.cuh"

#define GTAP_NUM_QUEUES 3

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
    #pragma gtap task queue((n - 1) < d_cutoff ? 1 : 0)
    x = fib(n - 1);
    #pragma gtap task queue((n - 2) < d_cutoff ? 1 : 0)
    y = fib(n - 2);
    #pragma gtap taskwait queue(2)
    return x + y;
}

__global__ void fib_kernel(int n) {
    #pragma gtap entry
    d_result = fib(n);
}
