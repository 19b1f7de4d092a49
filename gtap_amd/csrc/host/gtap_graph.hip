// gtap_graph.hip -- host-side generators of the benchmark inputs.
//
// R-MAT edges (SplitMix64 quadrant draws, one independent stream per edge, so the edge list
// does not depend on the thread count) and the CSR layout by a stable counting sort on the
// source vertex; the key generator of the reference's evaluation/util/gen_vector.c:8-18
// (LCG, value = INT32_MIN + (x mod 2^32), kept as the uint32 bit pattern).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "gtap_launch.hpp"

namespace {

inline uint64_t splitmix64_next(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

unsigned host_threads() {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    return std::min(t, 16u);
}

template <class F>
void parallel_for(int64_t n, F f) {
    const unsigned nt = host_threads();
    if (n < 1 << 16 || nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const int64_t b = (int64_t)t * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

// Edge e of an R-MAT graph with 2^scale vertices: s = seed ^ ((e+1)*0xD1B54A32D192ED03), then
// per level the top 32 bits r of the next SplitMix64 output choose quadrant (0,0) r<ta,
// (0,1) r<tab, (1,0) r<tabc, else (1,1); bits are appended most-significant first.
extern "C" int gtap_rmat_edges(int scale, int64_t nedges, uint64_t seed, uint32_t ta, uint32_t tab, uint32_t tabc,
                               int32_t* src, int32_t* dst) {
    if (scale < 1 || scale > 30 || nedges < 0 || !src || !dst) return gtap::host::invalid_argument("gtap_rmat_edges");
    parallel_for(nedges, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            uint64_t s = seed ^ ((uint64_t)(i + 1) * 0xD1B54A32D192ED03ULL);
            uint32_t u = 0, v = 0;
            for (int l = 0; l < scale; ++l) {
                const uint32_t r = (uint32_t)(splitmix64_next(s) >> 32);
                const uint32_t qi = (r < ta) ? 0u : (r < tab ? 0u : (r < tabc ? 1u : 1u));
                const uint32_t qj = (r < ta) ? 0u : (r < tab ? 1u : (r < tabc ? 0u : 1u));
                u = (u << 1) | qi;
                v = (v << 1) | qj;
            }
            src[i] = (int32_t)u;
            dst[i] = (int32_t)v;
        }
    });
    return 0;
}

// CSR by a stable counting sort on the source vertex (row: nv+1 offsets, col: ne columns).
extern "C" int gtap_csr_from_edges(int32_t nv, int64_t ne, const int32_t* src, const int32_t* dst, int32_t* row,
                                   int32_t* col) {
    if (nv <= 0 || ne < 0 || ne > INT32_MAX || !row || (ne && (!src || !dst || !col)))
        return gtap::host::invalid_argument("gtap_csr_from_edges");
    std::vector<int64_t> cnt((size_t)nv + 1, 0);
    for (int64_t e = 0; e < ne; ++e) {
        if ((uint32_t)src[e] >= (uint32_t)nv) return gtap::host::invalid_argument("gtap_csr_from_edges: src out of range");
        cnt[(size_t)src[e] + 1] += 1;
    }
    for (int32_t v = 0; v < nv; ++v) cnt[(size_t)v + 1] += cnt[(size_t)v];
    for (int32_t v = 0; v <= nv; ++v) row[v] = (int32_t)cnt[(size_t)v];
    for (int64_t e = 0; e < ne; ++e) col[cnt[(size_t)src[e]]++] = dst[e];
    return 0;
}

// R-MAT graph straight to CSR (row: 2^scale + 1, col: 2^scale * edge_factor).
extern "C" int gtap_rmat_csr(int scale, int edge_factor, uint64_t seed, uint32_t ta, uint32_t tab, uint32_t tabc,
                             int32_t* row, int32_t* col) {
    if (scale < 1 || scale > 30 || edge_factor < 1) return gtap::host::invalid_argument("gtap_rmat_csr");
    const int64_t nv = (int64_t)1 << scale;
    const int64_t ne = nv * edge_factor;
    std::vector<int32_t> src((size_t)ne), dst((size_t)ne);
    int st = gtap_rmat_edges(scale, ne, seed, ta, tab, tabc, src.data(), dst.data());
    if (st) return st;
    return gtap_csr_from_edges((int32_t)nv, ne, src.data(), dst.data(), row, col);
}

// Keys of evaluation/util/gen_vector.c:8-18,157-162 as uint32 bit patterns.
extern "C" int gtap_gen_keys(uint64_t seed, uint64_t n, uint32_t* out) {
    if (!out && n) return gtap::host::invalid_argument("gtap_gen_keys");
    uint64_t x = seed;
    for (uint64_t i = 0; i < n; ++i) {
        x = x * 1103515245ULL + 12345ULL;
        out[i] = 0x80000000u + (uint32_t)x;
    }
    return 0;
}
