// gtap_io.hip -- the reference programs' on-disk input formats (host code only).
//
//   graph CSR  (examples/asynchronous_bfs/asynchronous_bfs_thread.cu:69-84):
//              int32 N, int32 M, int32 row[N+1], int32 col[M]
//   key vector (evaluation/util/gen_vector.c:21-60, save_array / load_array):
//              size_t n (8 bytes), int32 data[n]
//   SpMV CSR   (examples/spmv/spmv_thread.cu:70-86):
//              int32 rows, int32 cols, int32 nnz, int32 row_ptr[rows+1], int32 col_idx[nnz],
//              float64 val[nnz]
//
// All little-endian, no padding. Readers take caller-sized buffers: call the *_info function
// first for the sizes. A file shorter than its header says, or with negative sizes, is an
// invalid argument; nothing is read past what the header declares.
#include <cstdio>
#include <cstring>

#include "gtap_launch.hpp"

namespace {

struct File {
    FILE* f = nullptr;
    explicit File(const char* path, const char* mode) : f(path ? fopen(path, mode) : nullptr) {}
    ~File() {
        if (f) fclose(f);
    }
    bool read(void* dst, size_t bytes) { return bytes == 0 || fread(dst, 1, bytes, f) == bytes; }
    bool write(const void* src, size_t bytes) { return bytes == 0 || fwrite(src, 1, bytes, f) == bytes; }
    int64_t size() {
        const long at = ftell(f);
        if (fseek(f, 0, SEEK_END) != 0) return -1;
        const long end = ftell(f);
        fseek(f, at, SEEK_SET);
        return (int64_t)end;
    }
};

int bad(const char* what) { return gtap::host::invalid_argument(what); }

}  // namespace

extern "C" {

int gtap_csr_file_info(const char* path, int32_t* nv, int64_t* ne) {
    File F(path, "rb");
    if (!F.f) return bad("gtap_csr_file_info: cannot open file");
    int32_t h[2];
    if (!F.read(h, sizeof(h)) || h[0] < 0 || h[1] < 0) return bad("gtap_csr_file_info: bad header");
    if (F.size() < 8 + 4 * ((int64_t)h[0] + 1) + 4 * (int64_t)h[1]) return bad("gtap_csr_file_info: file is truncated");
    if (nv) *nv = h[0];
    if (ne) *ne = h[1];
    return 0;
}

int gtap_csr_file_read(const char* path, int32_t* row, int32_t* col) {
    int32_t nv = 0;
    int64_t ne = 0;
    if (int st = gtap_csr_file_info(path, &nv, &ne)) return st;
    if (!row || (ne > 0 && !col)) return bad("gtap_csr_file_read: null buffer");
    File F(path, "rb");
    if (!F.f || fseek(F.f, 8, SEEK_SET) != 0) return bad("gtap_csr_file_read: cannot open file");
    if (!F.read(row, 4 * ((size_t)nv + 1)) || !F.read(col, 4 * (size_t)ne)) return bad("gtap_csr_file_read: short read");
    if (int st = gtap::host::check_csr_rows(row, nv, ne, "gtap_csr_file_read")) return st;
    for (int64_t i = 0; i < ne; ++i)
        if (col[i] < 0 || col[i] >= nv) return bad("gtap_csr_file_read: column index out of range");
    return 0;
}

int gtap_csr_file_write(const char* path, int32_t nv, int64_t ne, const int32_t* row, const int32_t* col) {
    if (nv < 0 || ne < 0 || ne > 0x7FFFFFFF || !row || (ne > 0 && !col)) return bad("gtap_csr_file_write: bad arguments");
    File F(path, "wb");
    if (!F.f) return bad("gtap_csr_file_write: cannot create file");
    const int32_t h[2] = {nv, (int32_t)ne};
    if (!F.write(h, sizeof(h)) || !F.write(row, 4 * ((size_t)nv + 1)) || !F.write(col, 4 * (size_t)ne))
        return bad("gtap_csr_file_write: short write");
    return 0;
}

int gtap_keys_file_info(const char* path, uint64_t* n) {
    File F(path, "rb");
    if (!F.f) return bad("gtap_keys_file_info: cannot open file");
    uint64_t h = 0;
    if (!F.read(&h, sizeof(h)) || h > ((uint64_t)1 << 40)) return bad("gtap_keys_file_info: bad header");
    if (F.size() < 8 + 4 * (int64_t)h) return bad("gtap_keys_file_info: file is truncated");
    if (n) *n = h;
    return 0;
}

// The int32 values of the file, as their 32-bit patterns.
int gtap_keys_file_read(const char* path, uint32_t* keys) {
    uint64_t n = 0;
    if (int st = gtap_keys_file_info(path, &n)) return st;
    if (n > 0 && !keys) return bad("gtap_keys_file_read: null buffer");
    File F(path, "rb");
    if (!F.f || fseek(F.f, 8, SEEK_SET) != 0) return bad("gtap_keys_file_read: cannot open file");
    if (!F.read(keys, 4 * (size_t)n)) return bad("gtap_keys_file_read: short read");
    return 0;
}

int gtap_keys_file_write(const char* path, uint64_t n, const uint32_t* keys) {
    if (n > 0 && !keys) return bad("gtap_keys_file_write: null buffer");
    File F(path, "wb");
    if (!F.f) return bad("gtap_keys_file_write: cannot create file");
    if (!F.write(&n, sizeof(n)) || !F.write(keys, 4 * (size_t)n)) return bad("gtap_keys_file_write: short write");
    return 0;
}

int gtap_spmv_file_info(const char* path, int32_t* rows, int32_t* cols, int32_t* nnz) {
    File F(path, "rb");
    if (!F.f) return bad("gtap_spmv_file_info: cannot open file");
    int32_t h[3];
    if (!F.read(h, sizeof(h)) || h[0] < 0 || h[1] < 0 || h[2] < 0) return bad("gtap_spmv_file_info: bad header");
    if (F.size() < 12 + 4 * ((int64_t)h[0] + 1) + 12 * (int64_t)h[2]) return bad("gtap_spmv_file_info: file is truncated");
    if (rows) *rows = h[0];
    if (cols) *cols = h[1];
    if (nnz) *nnz = h[2];
    return 0;
}

int gtap_spmv_file_read(const char* path, int32_t* row_ptr, int32_t* col_idx, double* val) {
    int32_t rows = 0, cols = 0, nnz = 0;
    if (int st = gtap_spmv_file_info(path, &rows, &cols, &nnz)) return st;
    if (!row_ptr || (nnz > 0 && (!col_idx || !val))) return bad("gtap_spmv_file_read: null buffer");
    File F(path, "rb");
    if (!F.f || fseek(F.f, 12, SEEK_SET) != 0) return bad("gtap_spmv_file_read: cannot open file");
    if (!F.read(row_ptr, 4 * ((size_t)rows + 1)) || !F.read(col_idx, 4 * (size_t)nnz) || !F.read(val, 8 * (size_t)nnz))
        return bad("gtap_spmv_file_read: short read");
    if (int st = gtap::host::check_csr_rows(row_ptr, rows, nnz, "gtap_spmv_file_read")) return st;
    for (int64_t i = 0; i < nnz; ++i)
        if (col_idx[i] < 0 || col_idx[i] >= cols) return bad("gtap_spmv_file_read: column index out of range");
    return 0;
}

int gtap_spmv_file_write(const char* path, int32_t rows, int32_t cols, int32_t nnz, const int32_t* row_ptr,
                         const int32_t* col_idx, const double* val) {
    if (rows < 0 || cols < 0 || nnz < 0 || !row_ptr || (nnz > 0 && (!col_idx || !val)))
        return bad("gtap_spmv_file_write: bad arguments");
    File F(path, "wb");
    if (!F.f) return bad("gtap_spmv_file_write: cannot create file");
    const int32_t h[3] = {rows, cols, nnz};
    if (!F.write(h, sizeof(h)) || !F.write(row_ptr, 4 * ((size_t)rows + 1)) || !F.write(col_idx, 4 * (size_t)nnz) ||
        !F.write(val, 8 * (size_t)nnz))
        return bad("gtap_spmv_file_write: short write");
    return 0;
}

}  // extern "C"
