// One task that spawns n detached children in one round, each adding to two device counters:
// with a small deque capacity (gtap_config.deque_capacity = 4096) the spawning wave publishes
// more than half its deque at the round's end, so it hands its oldest tasks to the overflow
// ring, from which idle waves take them (the overflow-ring tests).
#include <cuda_runtime.h>
#include "gtap_thread.cuh"

__device__ unsigned long long d_count;
__device__ unsigned long long d_sum;

#pragma gtap function
__device__ void leaf(int i) {
    atomicAdd(&d_count, 1ull);
    atomicAdd(&d_sum, (unsigned long long)i);
}

#pragma gtap function
__device__ void burst(int n) {
    for (int i = 0; i < n; ++i) {
        #pragma gtap task
        leaf(i);
    }
}

__global__ void burst_kernel(int n) {
    #pragma gtap entry
    burst(n);
}
