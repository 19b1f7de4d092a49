// Programs that misuse the task model on purpose, for the error-path tests: each makes the
// runtime record an error and end the launch normally.
//   chain(n, q, w): a chain of n tasks; the child goes to queue q, the parent resumes on w
//   (queue indices must be below GTAP_NUM_QUEUES);
//   leave(n): at n == 2 a task returns before the taskwait that joins its child;
//   wide(n): one task joins n children, more than a deque holds when n is large (joined
//   children must stay private until the round ends, so they cannot go to the overflow ring).
#include <cuda_runtime.h>
#include "gtap_thread.cuh"

#define GTAP_NUM_QUEUES 2

__device__ long long d_out;

#pragma gtap function
__device__ long long chain(int n, int q, int w) {
    if (n <= 0) return 1;
    long long a;
    #pragma gtap task queue(q)
    a = chain(n - 1, q, w);
    #pragma gtap taskwait queue(w)
    return a + 1;
}

#pragma gtap function
__device__ long long leave(int n) {
    if (n <= 0) return 1;
    long long a = 0;
    #pragma gtap task
    a = leave(n - 1);
    if (n == 2) return 7;
    #pragma gtap taskwait
    return a + 1;
}

#pragma gtap function
__device__ long long one(int i) {
    return i >= 0 ? 1 : 0;
}

#pragma gtap function
__device__ long long wide(int n) {
    for (int i = 0; i < n; ++i) {
        #pragma gtap task
        one(i);
    }
    #pragma gtap taskwait
    return n;
}

__global__ void chain_kernel(int n, int q, int w) {
    #pragma gtap entry
    d_out = chain(n, q, w);
}

__global__ void leave_kernel(int n) {
    #pragma gtap entry
    d_out = leave(n);
}

__global__ void wide_kernel(int n) {
    #pragma gtap entry
    d_out = wide(n);
}
