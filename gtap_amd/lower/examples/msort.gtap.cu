// Merge sort in the #pragma gtap dialect: two half-sorts as tasks, a taskwait, then a
// sequential merge. Keys written by one task and read by another go through
// gtap_coherent_load / gtap_coherent_store.
#include <cuda_runtime.h>
#include "gtap_thread.cuh"

__device__ void insertion_sort(unsigned* a, long n) {
    for (long i = 1; i < n; ++i) {
        unsigned v = gtap_coherent_load(a + i);
        long j = i - 1;
        while (j >= 0 && gtap_coherent_load(a + j) > v) {
            gtap_coherent_store(a + j + 1, gtap_coherent_load(a + j));
            --j;
        }
        gtap_coherent_store(a + j + 1, v);
    }
}

__device__ void merge_into(const unsigned* a, long na, const unsigned* b, long nb, unsigned* out) {
    long i = 0, j = 0, k = 0;
    while (i < na && j < nb) {
        unsigned x = gtap_coherent_load(a + i), y = gtap_coherent_load(b + j);
        if (y < x) { gtap_coherent_store(out + k, y); ++j; }
        else { gtap_coherent_store(out + k, x); ++i; }
        ++k;
    }
    while (i < na) { gtap_coherent_store(out + k, gtap_coherent_load(a + i)); ++i; ++k; }
    while (j < nb) { gtap_coherent_store(out + k, gtap_coherent_load(b + j)); ++j; ++k; }
}

#pragma gtap function
__device__ void msort(unsigned* a, unsigned* tmp, long n) {
    if (n <= 32) {
        insertion_sort(a, n);
        return;
    }
    long h = n / 2;
    #pragma gtap task
    msort(a, tmp, h);
    #pragma gtap task
    msort(a + h, tmp + h, n - h);
    #pragma gtap taskwait
    merge_into(a, h, a + h, n - h, tmp);
    for (long i = 0; i < n; ++i) gtap_coherent_store(a + i, gtap_coherent_load(tmp + i));
}

__global__ void msort_kernel(unsigned* a, unsigned* tmp, long n) {
    #pragma gtap entry
    msort(a, tmp, n);
}
