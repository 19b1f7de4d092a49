// Asynchronous BFS in the #pragma gtap dialect over a CSR graph bound to __device__
// pointers: the kernel sets the source's depth before the entry, a task lowers its
// neighbours' depths with atomicMin and spawns a task for each one it lowered (detached:
// no taskwait, the launch ends when every task has run). Long adjacency lists split into
// ranges of SPLIT edges.
#include <cuda_runtime.h>
#include "gtap_thread.cuh"

#define SPLIT 32

__device__ int* g_row;
__device__ int* g_col;
__device__ int* g_depth;

#pragma gtap function
__device__ void visit(int v, int s, int e) {
    if (s < 0) {
        s = g_row[v];
        e = g_row[v + 1];
        if (e - s > SPLIT) {
            for (int b = s; b < e; b += SPLIT) {
                int f = b + SPLIT < e ? b + SPLIT : e;
                #pragma gtap task
                visit(v, b, f);
            }
            return;
        }
    }
    int d = load_L2(&g_depth[v]) + 1;
    for (int i = s; i < e; ++i) {
        int u = g_col[i];
        if (atomicMin(&g_depth[u], d) > d) {
            #pragma gtap task
            visit(u, -1, -1);
        }
    }
}

__global__ void bfs_kernel(int source) {
    g_depth[source] = 0;
    #pragma gtap entry
    visit(source, -1, -1);
}
