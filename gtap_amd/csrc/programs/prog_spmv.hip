// prog_spmv.hip -- divide-and-conquer sparse matrix-vector product y = A x (CSR, float64) as
// a thread-executed task program.
//
// Source program (the task shape of the reference's examples/spmv/spmv_thread.cu:34-66):
//
//   #pragma gtap function
//   __device__ void spmv_row(int i) {
//       double sum = 0.0;
//       for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) sum = fma(val[k], x[col_idx[k]], sum);
//       y[i] = sum;
//   }
//   #pragma gtap function
//   __device__ void range_spmv(int begin, int end) {
//       if (end - begin <= CUTOFF) {
//           for (int i = begin; i < end; ++i) {
//               #pragma gtap task
//               spmv_row(i);
//           }
//       } else {
//           int mid = (begin + end) >> 1;
//           #pragma gtap task
//           range_spmv(begin, mid);
//           #pragma gtap task
//           range_spmv(mid, end);
//       }
//   }
//
// All spawns are detached (no taskwait); the launch ends by quiescence. The row sum is an
// explicit fused multiply-add chain in column order, so the result is the same bit pattern
// as a sequential fma loop over the row on any device (the reference writes `sum += v * x`,
// which a compiler may or may not contract).
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "../device/gtap_abi.hpp"
#include "../host/gtap_launch.hpp"

namespace gtap {
namespace spmv {

enum : uint32_t { F_RANGE = 1, F_ROW = 2 };
struct RangeArg {
    int32_t begin;
    int32_t end;
};
struct SpmvArgs {
    const int32_t* row_ptr;
    const int32_t* col_idx;
    const double* val;
    const double* x;
    double* y;
    int32_t rows;
    int32_t cols;
    int32_t cutoff;
    int32_t nnz;  // row_ptr[rows]: a row's offsets must satisfy 0 <= row_ptr[i] <= row_ptr[i+1] <= nnz
};

struct Program {
    static constexpr int kQueues = 1;
    static constexpr bool kNoTaskwait = true;  // detached spawns only
    static constexpr int kFrameWords = 2;      // begin, end (a row task uses begin)
    static constexpr uint32_t kMaxFunc = F_ROW;
    static constexpr uint32_t kFrameBytes = 16;
    using Args = SpmvArgs;
    template <class L>
    __device__ __forceinline__ static void run(L& lane);
};

// One row: the column indices and values of 8 entries are loaded together, then their x
// entries, then the fma chain runs in column order.
template <class L>
__device__ __forceinline__ void row_task(L& lane) {
    const SpmvArgs& g = *lane.args;
    const int32_t i = lane.template arg<int32_t>(0);
    if (i < 0 || i >= g.rows) {
        GTAP_DEVICE_ERROR(*lane.dev, GTAP_ERROR_INVALID_ARGUMENT, lane.tid, -1, i, g.rows);
        __gtap_finish_task(lane);
        return;
    }
    const int32_t kb = g.row_ptr[i], ke = g.row_ptr[i + 1];
    if (kb < 0 || ke < kb || ke > g.nnz) {  // malformed offsets: never read past col_idx / val
        GTAP_DEVICE_ERROR(*lane.dev, GTAP_ERROR_INVALID_ARGUMENT, lane.tid, -1, i, g.nnz);
        __gtap_finish_task(lane);
        return;
    }
    double sum = 0.0;
    constexpr int kChunk = 8;
    for (int32_t b = kb; b < ke; b += kChunk) {
        int32_t c[kChunk];
        double v[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            c[j] = (b + j < ke) ? g.col_idx[b + j] : 0;
            v[j] = (b + j < ke) ? g.val[b + j] : 0.0;
        }
        bool bad = false;
#pragma unroll
        for (int j = 0; j < kChunk; ++j) bad |= (b + j < ke) && (c[j] < 0 || c[j] >= g.cols);
        if (bad) {
            GTAP_DEVICE_ERROR(*lane.dev, GTAP_ERROR_INVALID_ARGUMENT, lane.tid, -1, i, g.cols);
            break;
        }
        double xv[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) xv[j] = (b + j < ke) ? g.x[c[j]] : 0.0;
#pragma unroll
        for (int j = 0; j < kChunk; ++j)
            if (b + j < ke) sum = __builtin_fma(v[j], xv[j], sum);
    }
    g.y[i] = sum;
    __gtap_finish_task(lane);
}

template <class L>
__device__ __forceinline__ void range_task(L& lane) {
    const SpmvArgs& g = *lane.args;
    const int32_t begin = lane.template arg<int32_t>(0);
    const int32_t end = lane.template arg<int32_t>(4);
    if (end - begin <= g.cutoff) {
        for (int32_t i = begin; i < end; ++i) __gtap_spawn_task_detached(lane, F_ROW, 0, RangeArg{i, 0});
    } else {
        const int32_t mid = (begin + end) >> 1;
        __gtap_spawn_task_detached(lane, F_RANGE, 0, RangeArg{begin, mid});
        __gtap_spawn_task_detached(lane, F_RANGE, 0, RangeArg{mid, end});
    }
    __gtap_finish_task(lane);
}

template <class L>
__device__ __forceinline__ void Program::run(L& lane) {
    if (lane.func == F_ROW) row_task(lane);
    else range_task(lane);
}

}  // namespace spmv
}  // namespace gtap

using SpmvProgram = gtap::spmv::Program;
static_assert(SpmvProgram::kFrameBytes + gtap::kHdrBytes <= gtap::host::kBuiltinStride, "frame exceeds stride");

// y = A x for a CSR matrix already in device memory (row_ptr: rows+1 int32, col_idx: nnz
// int32, val: nnz float64, x: cols float64, y: rows float64). Rows are split in halves down to
// ranges of at most `cutoff` rows (0 -> 16), whose rows become one task each.
// Replaces the program of examples/spmv/spmv_thread.cu:34-66.
extern "C" int gtap_spmv(const int32_t* row_ptr, const int32_t* col_idx, const double* val, const double* x,
                         double* y, int32_t rows, int32_t cols, int32_t cutoff, gtap_run_stats* stats) {
    using namespace gtap;
    if (!row_ptr || rows < 0 || cols < 0 || (rows > 0 && !y) || (cols > 0 && !x))
        return host::invalid_argument("gtap_spmv: bad arguments");
    if (cutoff <= 0) cutoff = 16;
    if (rows == 0) {
        if (stats) memset(stats, 0, sizeof(*stats));
        return 0;
    }
    int32_t nnz = 0;  // the entry count is row_ptr[rows]; every row range is checked against it
    GTAP_HIP_TRY(hipMemcpy(&nnz, row_ptr + rows, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (nnz < 0) return host::invalid_argument("gtap_spmv: negative entry count");
    if (nnz > 0 && (!col_idx || !val)) return host::invalid_argument("gtap_spmv: bad arguments");
    spmv::RangeArg root{0, rows};
    gtap_job job = host::make_job(spmv::F_RANGE, 0, &root, sizeof(root));
    spmv::SpmvArgs a{row_ptr, col_idx, val, x, y, rows, cols, cutoff, nnz};
    return host::run_single_job<SpmvProgram>(a, job, nullptr, stats, GTAP_TERM_QUIESCE);
}

// Host-buffer convenience: copies the matrix and x in, runs gtap_spmv, copies y out.
extern "C" int gtap_spmv_host(const int32_t* row_ptr, const int32_t* col_idx, const double* val, const double* x,
                              double* y, int32_t rows, int32_t cols, int32_t cutoff, gtap_run_stats* stats) {
    using namespace gtap;
    if (!row_ptr || rows < 0 || cols < 0 || (rows > 0 && !y) || (cols > 0 && !x))
        return host::invalid_argument("gtap_spmv_host: bad arguments");
    if (rows == 0) return gtap_spmv(row_ptr, col_idx, val, x, y, 0, cols, cutoff, stats);
    const int64_t nnz = row_ptr[rows];
    if (nnz < 0) return host::invalid_argument("gtap_spmv_host: negative entry count");
    if (int st = host::check_csr_rows(row_ptr, rows, nnz, "gtap_spmv_host")) return st;
    if (nnz > 0 && (!col_idx || !val)) return host::invalid_argument("gtap_spmv_host: bad arguments");
    int32_t *drp = nullptr, *dci = nullptr;
    double *dv = nullptr, *dx = nullptr, *dy = nullptr;
    int st = 0;
    hipError_t e = hipMalloc(&drp, (size_t)(rows + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&dci, (size_t)(nnz > 0 ? nnz : 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&dv, (size_t)(nnz > 0 ? nnz : 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&dx, (size_t)(cols > 0 ? cols : 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&dy, (size_t)rows * 8);
    if (e == hipSuccess) e = hipMemcpy(drp, row_ptr, (size_t)(rows + 1) * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz > 0) e = hipMemcpy(dci, col_idx, (size_t)nnz * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz > 0) e = hipMemcpy(dv, val, (size_t)nnz * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && cols > 0) e = hipMemcpy(dx, x, (size_t)cols * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) st = host::hip_status(e, "gtap_spmv_host copy-in");
    if (!st) st = gtap_spmv(drp, dci, dv, dx, dy, rows, cols, cutoff, stats);
    if (!st) {
        e = hipMemcpy(y, dy, (size_t)rows * 8, hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = host::hip_status(e, "gtap_spmv_host copy-out");
    }
    (void)hipFree(drp);
    (void)hipFree(dci);
    (void)hipFree(dv);
    (void)hipFree(dx);
    (void)hipFree(dy);
    return st;
}

namespace {

uint64_t spmv_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

// A random CSR matrix of the shape of examples/spmv/spmv_thread.cu:89-107 (rows x rows;
// row i has 1 + (h mod (2*avg_nnz - 1)) column draws, sorted and deduplicated; values uniform
// in [-1, 1)), from a splitmix64 stream keyed by (seed, row, draw), so every row can be
// generated on its own: rows are generated by several host threads, twice (counts, then
// entries). Two calls: with row_ptr only (col_idx = val = NULL) it fills row_ptr and
// *nnz_out; with all buffers it fills the matrix.
extern "C" int gtap_spmv_random(int32_t rows, int32_t avg_nnz, uint64_t seed, int32_t* row_ptr, int32_t* col_idx,
                                double* val, int64_t* nnz_out) {
    if (rows < 0 || avg_nnz < 1 || avg_nnz > (1 << 24) || !row_ptr) return gtap::host::invalid_argument("gtap_spmv_random");
    const uint64_t span = 2ull * (uint64_t)avg_nnz - 1ull;
    // the sorted, deduplicated columns of row i in `cols`; returns the row's key
    auto row_cols = [&](int32_t i, std::vector<int32_t>& cols) {
        const uint64_t key = spmv_mix(seed ^ (0x9E3779B97F4A7C15ULL * ((uint64_t)i + 1)));
        const uint32_t cnt = 1u + (uint32_t)(key % span);
        cols.resize(cnt);
        for (uint32_t k = 0; k < cnt; ++k) cols[k] = (int32_t)(spmv_mix(key + 2ull * k + 1ull) % (uint64_t)rows);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        return key;
    };
    unsigned nth = std::thread::hardware_concurrency();
    nth = nth == 0 ? 1u : (nth > 32u ? 32u : nth);
    if ((int64_t)rows < 4096) nth = 1;
    auto parallel_rows = [&](auto&& body) {
        std::vector<std::thread> pool;
        const int64_t per = ((int64_t)rows + nth - 1) / nth;
        for (unsigned t = 0; t < nth; ++t) {
            const int32_t r0 = (int32_t)std::min<int64_t>((int64_t)t * per, rows);
            const int32_t r1 = (int32_t)std::min<int64_t>((int64_t)(t + 1) * per, rows);
            pool.emplace_back([&body, r0, r1] {
                std::vector<int32_t> cols;
                for (int32_t i = r0; i < r1; ++i) body(i, cols);
            });
        }
        for (auto& th : pool) th.join();
    };
    std::vector<int32_t> count((size_t)rows);
    parallel_rows([&](int32_t i, std::vector<int32_t>& cols) {
        row_cols(i, cols);
        count[(size_t)i] = (int32_t)cols.size();
    });
    int64_t nnz = 0;
    row_ptr[0] = 0;
    for (int32_t i = 0; i < rows; ++i) {
        nnz += count[(size_t)i];
        if (nnz > 0x7FFFFFFF) return gtap::host::invalid_argument("gtap_spmv_random: more than 2^31-1 entries");
        row_ptr[i + 1] = (int32_t)nnz;
    }
    if (col_idx && val) {
        parallel_rows([&](int32_t i, std::vector<int32_t>& cols) {
            const uint64_t key = row_cols(i, cols);
            const int64_t at = row_ptr[i];
            for (size_t k = 0; k < cols.size(); ++k) {
                col_idx[at + (int64_t)k] = cols[k];
                const uint64_t r = spmv_mix(key + 2ull * (uint64_t)cols[k] + 2ull);
                val[at + (int64_t)k] = (double)(r >> 11) * (2.0 / 9007199254740992.0) - 1.0;
            }
        });
    }
    if (nnz_out) *nnz_out = nnz;
    return 0;
}
