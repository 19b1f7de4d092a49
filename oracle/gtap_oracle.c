/*
 * gtap_oracle.c -- CPU restatement of the GTaP hot-path programs (test infrastructure).
 *
 * This is the checker, never the product: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it. See gtap_oracle.h for the reference lines each
 * function follows and for how each one is pinned.
 */
#include "gtap_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- fib */

/* seq_fib.c:5-10 */
int64_t gtap_oracle_fib(int n) {
    if (n < 2) return n;
    return gtap_oracle_fib(n - 1) + gtap_oracle_fib(n - 2);
}

/* One task per call of gtap_fib.cu:10-22 (root included). */
int64_t gtap_oracle_fib_tasks(int n) {
    if (n < 2) return 1;
    return 1 + gtap_oracle_fib_tasks(n - 1) + gtap_oracle_fib_tasks(n - 2);
}

/* omp_fib.c:5-15, with a serial cutoff so the CPU baseline is not task-overhead bound
 * when cutoff > 1 (cutoff <= 1 reproduces omp_fib.c exactly). */
static int64_t fib_omp_rec(int n, int cutoff) {
    if (n < 2) return n;
    if (n < cutoff) return gtap_oracle_fib(n);
    int64_t a = 0, b = 0;
#pragma omp task shared(a)
    a = fib_omp_rec(n - 1, cutoff);
#pragma omp task shared(b)
    b = fib_omp_rec(n - 2, cutoff);
#pragma omp taskwait
    return a + b;
}

int64_t gtap_oracle_fib_omp(int n, int cutoff) {
    int64_t r = 0;
#pragma omp parallel
    {
#pragma omp single
        r = fib_omp_rec(n, cutoff);
    }
    return r;
}

int gtap_oracle_omp_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------- n-queens */

/* seq_nq.c:12-33; every call is one task of the no-cutoff task program. */
static void nq_rec(int n, int row, uint32_t col, uint32_t left, uint32_t right,
                   int64_t* sol, int64_t* tasks) {
    *tasks += 1;
    if (row == n) {
        *sol += 1;
        return;
    }
    uint32_t mask = (n < 32 ? (1u << n) - 1u : 0xFFFFFFFFu);
    uint32_t avail = mask & ~(col | left | right);
    while (avail) {
        uint32_t p = avail & (0u - avail);
        avail -= p;
        nq_rec(n, row + 1, col | p, (left | p) << 1, (right | p) >> 1, sol, tasks);
    }
}

void gtap_oracle_nq(int n, int64_t* solutions, int64_t* tasks) {
    int64_t s = 0, t = 0;
    nq_rec(n, 0, 0, 0, 0, &s, &t);
    *solutions = s;
    *tasks = t;
}

/* Nodes of the search tree that have at least one child (they suspend at the taskwait
 * and run a second time). */
static int64_t nq_internal_rec(int n, int row, uint32_t col, uint32_t left, uint32_t right) {
    if (row == n) return 0;
    uint32_t mask = (n < 32 ? (1u << n) - 1u : 0xFFFFFFFFu);
    uint32_t avail = mask & ~(col | left | right);
    if (!avail) return 0;
    int64_t c = 1;
    while (avail) {
        uint32_t p = avail & (0u - avail);
        avail -= p;
        c += nq_internal_rec(n, row + 1, col | p, (left | p) << 1, (right | p) >> 1);
    }
    return c;
}

int64_t gtap_oracle_nq_internal(int n) { return nq_internal_rec(n, 0, 0, 0, 0); }

static int64_t nq_count_serial(int n, int row, uint32_t col, uint32_t left, uint32_t right) {
    if (row == n) return 1;
    uint32_t mask = (n < 32 ? (1u << n) - 1u : 0xFFFFFFFFu);
    uint32_t avail = mask & ~(col | left | right);
    int64_t s = 0;
    while (avail) {
        uint32_t p = avail & (0u - avail);
        avail -= p;
        s += nq_count_serial(n, row + 1, col | p, (left | p) << 1, (right | p) >> 1);
    }
    return s;
}

/* omp_nq.c:23-43 shape (task per legal column above the cutoff row), but each task
 * returns its int64 count and the parent sums them after a taskwait. */
static int64_t nq_omp_rec(int n, int cutoff, int row, uint32_t col, uint32_t left, uint32_t right) {
    if (row > cutoff) return nq_count_serial(n, row, col, left, right);
    if (row == n) return 1;
    uint32_t mask = (n < 32 ? (1u << n) - 1u : 0xFFFFFFFFu);
    uint32_t avail = mask & ~(col | left | right);
    int64_t counts[32];
    int k = 0;
    while (avail) {
        uint32_t p = avail & (0u - avail);
        avail -= p;
        int64_t* slot = &counts[k++];
#pragma omp task firstprivate(p, slot)
        *slot = nq_omp_rec(n, cutoff, row + 1, col | p, (left | p) << 1, (right | p) >> 1);
    }
#pragma omp taskwait
    int64_t s = 0;
    for (int i = 0; i < k; ++i) s += counts[i];
    return s;
}

int64_t gtap_oracle_nq_omp(int n, int cutoff) {
    int64_t r = 0;
#pragma omp parallel
    {
#pragma omp single
        r = nq_omp_rec(n, cutoff, 0, 0, 0, 0);
    }
    return r;
}

/* ---------------------------------------------------------------- keys */

/* gen_vector.c:8-18 (rng_next / rng_int) and 157-162 (rng_int(INT_MIN, INT_MAX)).
 * rng_int computes min + (int)(x % 2^32); done here in uint32 arithmetic, which gives
 * the same bits without relying on signed overflow. */
void gtap_oracle_gen_keys_u32(uint64_t seed, size_t n, uint32_t* out) {
    uint64_t x = seed;
    for (size_t i = 0; i < n; ++i) {
        x = x * 1103515245ULL + 12345ULL;
        out[i] = 0x80000000u + (uint32_t)(x % 4294967296ULL);
    }
}

void gtap_oracle_gen_vector(uint64_t seed, size_t n, int32_t* out) {
    gtap_oracle_gen_keys_u32(seed, n, (uint32_t*)out);
}

/* ---------------------------------------------------------------- cilksort */

/* gtap_cilksort.cu:40-50 */
int64_t gtap_oracle_binary_search_u32(const uint32_t* x, int64_t len, uint32_t val) {
    int64_t low = -1, high = len;
    while (low + 1 < high) {
        int64_t mid = (low + high) / 2;
        if (x[mid] <= val) low = mid;
        else high = mid;
    }
    return low;
}

/* gtap_cilksort.cu:18-37 (takes from b on ties, as `a_val < b_val` does) */
static void merge_seq_u32(const uint32_t* a, int64_t na, const uint32_t* b, int64_t nb, uint32_t* dst) {
    int64_t i = 0, j = 0, t = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) dst[t++] = a[i++];
        else dst[t++] = b[j++];
    }
    while (i < na) dst[t++] = a[i++];
    while (j < nb) dst[t++] = b[j++];
}

/* gtap_cilksort.cu:90-127 */
static void cilkmerge_u32(const uint32_t* a, int64_t na, const uint32_t* b, int64_t nb, uint32_t* dst,
                          int64_t merge_cutoff) {
    if (na < nb) {
        const uint32_t* tp = a; a = b; b = tp;
        int64_t tn = na; na = nb; nb = tn;
    }
    if (nb == 0) {
        if (na > 0) memmove(dst, a, (size_t)na * sizeof(uint32_t));
        return;
    }
    if (na + nb < merge_cutoff) {
        merge_seq_u32(a, na, b, nb, dst);
        return;
    }
    int64_t a_split = (na + 1) / 2;
    int64_t b_split = gtap_oracle_binary_search_u32(b, nb, a[a_split - 1]) + 1;
    cilkmerge_u32(a, a_split, b, b_split, dst, merge_cutoff);
    cilkmerge_u32(a + a_split, na - a_split, b + b_split, nb - b_split, dst + a_split + b_split, merge_cutoff);
}

static void insertion_sort_u32(uint32_t* a, int64_t n) {
    for (int64_t i = 1; i < n; ++i) {
        uint32_t v = a[i];
        int64_t j = i - 1;
        while (j >= 0 && a[j] > v) {
            a[j + 1] = a[j];
            --j;
        }
        a[j + 1] = v;
    }
}

static int cmp_u32(const void* x, const void* y) {
    uint32_t a = *(const uint32_t*)x, b = *(const uint32_t*)y;
    return (a > b) - (a < b);
}

/* gtap_cilksort.cu:129-167 */
static void cilksort_rec_u32(uint32_t* arr, int64_t n, uint32_t* tmp, int64_t leaf_max, int64_t merge_cutoff) {
    if (n <= leaf_max) {
        if (n <= 64) insertion_sort_u32(arr, n);
        else qsort(arr, (size_t)n, sizeof(uint32_t), cmp_u32);
        return;
    }
    int64_t len12 = n / 2;
    int64_t len1 = len12 / 2;
    int64_t len2 = len12 - len1;
    int64_t len34 = n - len12;
    int64_t len3 = len34 / 2;
    int64_t len4 = len34 - len3;
    cilksort_rec_u32(arr, len1, tmp, leaf_max, merge_cutoff);
    cilksort_rec_u32(arr + len1, len2, tmp + len1, leaf_max, merge_cutoff);
    cilksort_rec_u32(arr + len12, len3, tmp + len12, leaf_max, merge_cutoff);
    cilksort_rec_u32(arr + len12 + len3, len4, tmp + len12 + len3, leaf_max, merge_cutoff);
    cilkmerge_u32(arr, len1, arr + len1, len2, tmp, merge_cutoff);
    cilkmerge_u32(arr + len12, len3, arr + len12 + len3, len4, tmp + len12, merge_cutoff);
    cilkmerge_u32(tmp, len12, tmp + len12, len34, arr, merge_cutoff);
}

void gtap_oracle_cilksort_u32(uint32_t* arr, uint32_t* tmp, int64_t n, int64_t leaf_max, int64_t merge_cutoff) {
    if (n <= 0) return;
    if (leaf_max < 1) leaf_max = 1;
    /* the binary split only shrinks a merge of >= 3 keys (a 1+1 merge would split into
     * itself), so merges below 3 keys are always sequential */
    if (merge_cutoff < 3) merge_cutoff = 3;
    cilksort_rec_u32(arr, n, tmp, leaf_max, merge_cutoff);
}

/* ---------------------------------------------------------------- mergesort */

/* seq_mergesort.cpp:38-50 */
static void merge_cpu_u32(uint32_t* a, uint32_t* buf, int64_t left, int64_t mid, int64_t right) {
    int64_t i = left, j = mid + 1, k = left;
    while (i <= mid && j <= right) {
        if (a[i] <= a[j]) buf[k++] = a[i++];
        else buf[k++] = a[j++];
    }
    while (i <= mid) buf[k++] = a[i++];
    while (j <= right) buf[k++] = a[j++];
    for (int64_t t = left; t <= right; ++t) a[t] = buf[t];
}

/* seq_mergesort.cpp:52-60 */
static void mergesort_rec_u32(uint32_t* a, uint32_t* buf, int64_t left, int64_t right) {
    if (left >= right) return;
    int64_t mid = left + ((right - left) >> 1);
    mergesort_rec_u32(a, buf, left, mid);
    mergesort_rec_u32(a, buf, mid + 1, right);
    merge_cpu_u32(a, buf, left, mid, right);
}

void gtap_oracle_mergesort_u32(uint32_t* a, uint32_t* buf, int64_t n) {
    if (n <= 0) return;
    mergesort_rec_u32(a, buf, 0, n - 1);
}

/* ---------------------------------------------------------------- R-MAT + BFS */

static uint64_t splitmix64_next(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void gtap_oracle_rmat_edges(int scale, int64_t nedges, uint64_t seed,
                            uint32_t ta, uint32_t tab, uint32_t tabc,
                            int32_t* src, int32_t* dst) {
    for (int64_t e = 0; e < nedges; ++e) {
        uint64_t s = seed ^ ((uint64_t)(e + 1) * 0xD1B54A32D192ED03ULL);
        uint32_t u = 0, v = 0;
        for (int l = 0; l < scale; ++l) {
            uint32_t r = (uint32_t)(splitmix64_next(&s) >> 32);
            uint32_t i, j;
            if (r < ta) { i = 0; j = 0; }
            else if (r < tab) { i = 0; j = 1; }
            else if (r < tabc) { i = 1; j = 0; }
            else { i = 1; j = 1; }
            u = (u << 1) | i;
            v = (v << 1) | j;
        }
        src[e] = (int32_t)u;
        dst[e] = (int32_t)v;
    }
}

void gtap_oracle_csr_from_edges(int32_t nv, int64_t ne, const int32_t* src, const int32_t* dst,
                                int32_t* row, int32_t* col) {
    int64_t* cnt = (int64_t*)calloc((size_t)nv + 1, sizeof(int64_t));
    for (int64_t e = 0; e < ne; ++e) cnt[src[e] + 1] += 1;
    for (int32_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
    for (int32_t v = 0; v <= nv; ++v) row[v] = (int32_t)cnt[v];
    for (int64_t e = 0; e < ne; ++e) col[cnt[src[e]]++] = dst[e];
    free(cnt);
}

/* asynchronous_bfs_thread.cu:145-174 */
void gtap_oracle_bfs(int32_t nv, const int32_t* row, const int32_t* col, int32_t source, int32_t* depth) {
    const int32_t INF = 0x3f3f3f3f;
    for (int32_t i = 0; i < nv; ++i) depth[i] = INF;
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv > 0 ? nv : 1));
    int64_t qh = 0, qt = 0;
    depth[source] = 0;
    q[qt++] = source;
    while (qh < qt) {
        int32_t v = q[qh++];
        int32_t dv = depth[v];
        for (int32_t e = row[v]; e < row[v + 1]; ++e) {
            int32_t u = col[e];
            if (dv + 1 < depth[u]) {
                depth[u] = dv + 1;
                q[qt++] = u;
            }
        }
    }
    free(q);
}

/* ---------------------------------------------------------------- job stream */

void gtap_oracle_jobs(uint64_t seed, int32_t njobs, int32_t* kind, int32_t* param, uint64_t* key_seed) {
    for (int32_t i = 0; i < njobs; ++i) {
        uint64_t s = seed ^ ((uint64_t)(2 * (int64_t)i + 1) * 0xD1B54A32D192ED03ULL);
        uint64_t r0 = splitmix64_next(&s);
        uint64_t r1 = splitmix64_next(&s);
        uint32_t k = (uint32_t)(r0 >> 32) % 10u;
        key_seed[i] = 0;
        if (k < 4) {
            kind[i] = 0;
            param[i] = 18 + (int32_t)((uint32_t)(r1 >> 32) % 11u);
        } else if (k < 7) {
            kind[i] = 1;
            param[i] = 8 + (int32_t)((uint32_t)(r1 >> 32) % 5u);
        } else {
            double u = (double)(r1 >> 11) * (1.0 / 9007199254740992.0);
            double len = floor(exp2(10.0 + 10.0 * u));
            if (len < 1024.0) len = 1024.0;
            if (len > 1048576.0) len = 1048576.0;
            kind[i] = 2;
            param[i] = (int32_t)len;
            key_seed[i] = r1 ^ 0x94D049BB133111EBULL;
        }
    }
}

uint64_t gtap_oracle_sort_checksum(const uint32_t* a, int64_t n) {
    uint64_t sum = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint64_t z = ((uint64_t)i << 32) ^ (uint64_t)a[i];
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        sum += z ^ (z >> 31);
    }
    return sum;
}

uint64_t gtap_oracle_hash_u32(const uint32_t* a, int64_t n) {
    uint64_t h = 0xcbf29ce484222325ULL;
    const unsigned char* p = (const unsigned char*)a;
    for (int64_t i = 0; i < n * 4; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

/* ------------------------------------------------------------------ SpMV, binary tree */

void gtap_oracle_spmv(int32_t rows, const int32_t* row_ptr, const int32_t* col_idx, const double* val,
                      const double* x, double* y) {
    for (int32_t i = 0; i < rows; ++i) {
        double sum = 0.0;
        for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) sum = fma(val[k], x[col_idx[k]], sum);
        y[i] = sum;
    }
}

static uint32_t oracle_xs32(uint32_t x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

static uint32_t oracle_sat_u32(double v) {
    if (!(v > 0.0)) return 0u;
    if (v >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

double gtap_oracle_tree_node(int32_t node, int32_t mem_ops, int32_t compute_iters, const double* input,
                             int64_t input_n) {
    const uint32_t seed = oracle_xs32((uint32_t)node ^ 0x9e3779b9u);
    const uint32_t mask = (uint32_t)input_n - 1u;
    double acc = 0.0;
    for (int32_t m = 0; m < mem_ops; ++m) acc += input[oracle_xs32(seed + (uint32_t)m * 747796405u) & mask];
    uint32_t mix = (0x9e3779b9u ^ (uint32_t)node) + oracle_sat_u32(acc);
    for (int32_t it = 0; it < compute_iters; ++it) {
        const double xv = (double)oracle_xs32(seed + (uint32_t)it * 747796405u);
        const double yv = fma(xv, 1.0000001192092896, 0.9999999403953552);
        uint64_t bits;
        memcpy(&bits, &yv, sizeof(bits));
        mix ^= (uint32_t)bits;
    }
    return (double)mix;
}

void gtap_oracle_tree(int32_t height, int32_t mem_ops, int32_t compute_iters, const double* input, int64_t input_n,
                      double* out) {
    const int64_t nodes = ((int64_t)1 << (height + 1)) - 1;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < nodes; ++v) out[v] = gtap_oracle_tree_node((int32_t)v, mem_ops, compute_iters, input, input_n);
}

/* ---------------------------------------------------------------- stream CPU baseline */

/* The mixed job stream on the CPU: every job of gtap_oracle_jobs(seed, njobs) computed by the
 * sequential programs (fib: seq_fib.c:5-10; n-queens: seq_nq.c:12-33; sort: the job's keys,
 * seq_mergesort.cpp:38-62 as gtap_oracle_mergesort_u32, then the sort checksum), one job per
 * loop iteration; `parallel` runs the loop as `#pragma omp parallel for schedule(dynamic)`
 * over the jobs (evaluation/2-comparison/mergesort/omp_mergesort.cpp:38-56 parallelizes
 * inside one sort; a stream of independent jobs parallelizes over the jobs). results[i]
 * receives job i's result; returns the wall time of the loop in seconds (omp_get_wtime). */
double gtap_oracle_stream_cpu(uint64_t seed, int32_t njobs, int32_t parallel, int64_t* results) {
    int32_t* kind = (int32_t*)malloc(sizeof(int32_t) * (size_t)(njobs > 0 ? njobs : 1));
    int32_t* param = (int32_t*)malloc(sizeof(int32_t) * (size_t)(njobs > 0 ? njobs : 1));
    uint64_t* kseed = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(njobs > 0 ? njobs : 1));
    gtap_oracle_jobs(seed, njobs, kind, param, kseed);
    const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic) if (parallel)
    for (int32_t i = 0; i < njobs; ++i) {
        int64_t r = 0;
        if (kind[i] == 0) {
            r = gtap_oracle_fib(param[i]);
        } else if (kind[i] == 1) {
            int64_t s = 0, t = 0;
            gtap_oracle_nq(param[i], &s, &t);
            r = s;
        } else {
            const int64_t n = param[i];
            uint32_t* a = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
            uint32_t* b = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
            gtap_oracle_gen_keys_u32(kseed[i], (size_t)n, a);
            gtap_oracle_mergesort_u32(a, b, n);
            r = (int64_t)gtap_oracle_sort_checksum(a, n);
            free(a);
            free(b);
        }
        results[i] = r;
    }
    const double dt = omp_get_wtime() - t0;
    free(kind);
    free(param);
    free(kseed);
    return dt;
}
