// A four-way merge sort over keys bound by the host to __device__ pointers, in the
// #pragma gtap dialect with plain array accesses: the lowering makes them coherent, and the
// split searches read through load_L2. Merges above a cutoff split in two at the median of
// the longer run (binary search of the other), as parallel merge tasks.
#include <cuda_runtime.h>
#include "gtap_thread.cuh"

#define SORT_CUTOFF 32
#define MERGE_CUTOFF 128

__device__ unsigned* g_keys;
__device__ unsigned* g_tmp;

__device__ void insertion(unsigned* a, int n) {
    for (int i = 1; i < n; ++i) {
        unsigned v = a[i];
        int j = i - 1;
        while (j >= 0 && a[j] > v) {
            a[j + 1] = a[j];
            --j;
        }
        a[j + 1] = v;
    }
}

// first index in x[0, n) whose key is greater than v
__device__ int upper(unsigned* x, int n, unsigned v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (load_L2(&x[mid]) <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

#pragma gtap function
__device__ void pmerge(unsigned* a, int na, unsigned* b, int nb, unsigned* out) {
    if (na < nb) {
        unsigned* t = a; a = b; b = t;
        int tn = na; na = nb; nb = tn;
    }
    if (na + nb <= MERGE_CUTOFF) {
        int i = 0, j = 0, k = 0;
        while (i < na && j < nb) {
            if (b[j] < a[i]) out[k++] = b[j++];
            else out[k++] = a[i++];
        }
        while (i < na) out[k++] = a[i++];
        while (j < nb) out[k++] = b[j++];
        return;
    }
    int ma = na / 2;
    int mb = upper(b, nb, a[ma]);
    #pragma gtap task
    pmerge(a, ma, b, mb, out);
    #pragma gtap task
    pmerge(a + ma, na - ma, b + mb, nb - mb, out + ma + mb);
    #pragma gtap taskwait
}

#pragma gtap function
__device__ void csort(unsigned* a, unsigned* t, int n) {
    if (n <= SORT_CUTOFF) {
        insertion(a, n);
        return;
    }
    int q = n / 4, r = n / 2, s = r + (n - r) / 2;
    #pragma gtap task
    csort(a, t, q);
    #pragma gtap task
    csort(a + q, t + q, r - q);
    #pragma gtap task
    csort(a + r, t + r, s - r);
    #pragma gtap task
    csort(a + s, t + s, n - s);
    #pragma gtap taskwait
    #pragma gtap task
    pmerge(a, q, a + q, r - q, t);
    #pragma gtap task
    pmerge(a + r, s - r, a + s, n - s, t + r);
    #pragma gtap taskwait
    #pragma gtap task
    pmerge(t, r, t + r, n - r, a);
    #pragma gtap taskwait
}

__global__ void csort_kernel(int n) {
    #pragma gtap entry
    csort(g_keys, g_tmp, n);
}
