// CSR transpose on the device (gfx950), for the backward structures of
// non-symmetric graphs.
//
// The reference builds no transpose: its layers reuse A for A^T
// (end2end_ori/gcn_final/mgat_csr/mgat_conv.py:57-74), which is exact only for
// symmetric A.  Steps, all on one stream:
//   1. count the entries of every column (integer atomics: order-free) and
//      write each entry's row id;
//   2. exclusive scan of the counts (rocPRIM) -> row pointers of A^T;
//   3. stable radix sort of (column, CSR position) pairs (rocPRIM): equal
//      columns keep CSR order, i.e. ascending source row, so A^T's rows come
//      out sorted and the result equals a stable host sort by (column, row);
//   4. gather: t_perm = CSR position of each A^T entry, t_col = its source row.
// Integer work throughout; the result is bit-exact and deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "attgnn.h"

namespace {

constexpr int THREADS = 256;

inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

struct TWs {
    int32_t *counts;     // [n_cols + 1]
    int32_t *row_of;     // [nnz]
    int32_t *keys_out;   // [nnz]
    int32_t *vals_out;   // [nnz]
    int32_t *status;     // [1]
    void *temp;          // rocPRIM scratch
    size_t temp_bytes;
};

size_t rocprim_temp_bytes(int64_t n_cols, int64_t nnz)
{
    // size queries only (no launch); 32 key bits bound every narrower sort
    size_t sort_bytes = 0, scan_bytes = 0;
    const hipError_t a = rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<const int32_t *>(nullptr),
                                                   static_cast<int32_t *>(nullptr),
                                                   rocprim::counting_iterator<int32_t>(0),
                                                   static_cast<int32_t *>(nullptr), size_t(nnz), 0, 32);
    const hipError_t b = rocprim::exclusive_scan(nullptr, scan_bytes, static_cast<const int32_t *>(nullptr),
                                                 static_cast<int32_t *>(nullptr), 0, size_t(n_cols + 1),
                                                 rocprim::plus<int32_t>());
    if (a != hipSuccess || b != hipSuccess) return size_t(1) << 40;   // refused by the size check
    return std::max(sort_bytes, scan_bytes);
}

int64_t ws_bytes(int64_t n_cols, int64_t nnz)
{
    return align256(4 * (n_cols + 1)) + 3 * align256(4 * nnz + 4) + align256(64) +
           align256(int64_t(rocprim_temp_bytes(n_cols, nnz)) + 256);
}

TWs ws_carve(void *base, int64_t n_cols, int64_t nnz)
{
    uint8_t *p = static_cast<uint8_t *>(base);
    TWs w;
    w.counts = reinterpret_cast<int32_t *>(p); p += align256(4 * (n_cols + 1));
    w.row_of = reinterpret_cast<int32_t *>(p); p += align256(4 * nnz + 4);
    w.keys_out = reinterpret_cast<int32_t *>(p); p += align256(4 * nnz + 4);
    w.vals_out = reinterpret_cast<int32_t *>(p); p += align256(4 * nnz + 4);
    w.status = reinterpret_cast<int32_t *>(p); p += align256(64);
    w.temp = p;
    w.temp_bytes = rocprim_temp_bytes(n_cols, nnz);
    return w;
}

inline unsigned grid_for(int64_t n, int64_t per_block, int64_t cap = 1 << 20)
{
    int64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return unsigned(g);
}

__global__ __launch_bounds__(THREADS) void zero_kernel(int32_t *p, int64_t n, int32_t *status)
{
    const int64_t t = int64_t(blockIdx.x) * THREADS + threadIdx.x;
    if (t == 0) status[0] = 0;
    for (int64_t i = t; i < n; i += int64_t(gridDim.x) * THREADS) p[i] = 0;
}

// One wave per row: row ids of its entries and the column counts.  Row ranges
// are clamped to [0, nnz) and out-of-range columns are skipped and flagged, so
// a malformed CSR cannot write out of bounds.
__global__ __launch_bounds__(THREADS) void count_kernel(const int32_t *rowptr, const int32_t *col, int64_t n_rows,
                                                        int64_t n_cols, int64_t nnz, TWs w)
{
    const int lane = threadIdx.x & 63;
    const int64_t nw = int64_t(gridDim.x) * (THREADS / 64);
    for (int64_t i = int64_t(blockIdx.x) * (THREADS / 64) + (threadIdx.x >> 6); i < n_rows; i += nw) {
        int64_t b = rowptr[i], e = rowptr[i + 1];
        if (b < 0 || e > nnz || e < b) {
            if (lane == 0) atomicOr(w.status, 1);
            continue;
        }
        for (int64_t m = b + lane; m < e; m += 64) {
            const int32_t c = col[m];
            w.row_of[m] = int32_t(i);
            if (c >= 0 && c < n_cols) atomicAdd(&w.counts[c], 1);
            else atomicOr(w.status, 1);
        }
    }
}

__global__ __launch_bounds__(THREADS) void gather_kernel(const TWs w, int64_t nnz, int32_t *t_col, int64_t *t_perm)
{
    for (int64_t q = int64_t(blockIdx.x) * THREADS + threadIdx.x; q < nnz; q += int64_t(gridDim.x) * THREADS) {
        const int32_t m = w.vals_out[q];
        t_perm[q] = m;
        t_col[q] = w.row_of[m];
    }
}

__global__ __launch_bounds__(THREADS) void status_kernel(const int32_t *rowptr, int64_t n_rows, int64_t nnz, TWs w,
                                                         int64_t *status_out)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const bool bad = w.status[0] != 0 || rowptr[0] != 0 || int64_t(rowptr[n_rows]) != nnz;
        status_out[0] = bad ? ATTGNN_E_FORMAT : ATTGNN_OK;
    }
}

// out[s] = map[s] < 0 ? -1 : remap[map[s]]
__global__ __launch_bounds__(THREADS) void remap_kernel(const int32_t *map, int64_t n, const int64_t *remap,
                                                        int64_t remap_n, int32_t *out)
{
    for (int64_t s = int64_t(blockIdx.x) * THREADS + threadIdx.x; s < n; s += int64_t(gridDim.x) * THREADS) {
        const int32_t m = map[s];
        out[s] = (m >= 0 && m < remap_n) ? int32_t(remap[m]) : -1;
    }
}

int bits_for(int64_t n)
{
    int b = 1;
    while ((int64_t(1) << b) < n) ++b;
    return b;
}

}  // namespace

extern "C" int64_t attgnn_csr_transpose_workspace_bytes(int64_t n_rows, int64_t n_cols, int64_t nnz)
{
    if (n_rows < 0 || n_cols < 0 || nnz < 0 || nnz > INT32_MAX || n_cols >= INT32_MAX) return -1;
    return ws_bytes(n_cols, nnz);
}

extern "C" int attgnn_csr_transpose_device(const int32_t *rowptr, const int32_t *col, int64_t n_rows, int64_t n_cols,
                                           int64_t nnz, int32_t *t_rowptr, int32_t *t_col, int64_t *t_perm,
                                           int64_t *status, void *workspace, int64_t workspace_bytes,
                                           attgnn_stream_t stream)
{
    if (n_rows < 0 || n_cols < 0 || nnz < 0 || nnz > INT32_MAX || n_cols >= INT32_MAX) return ATTGNN_E_ARG;
    if (!rowptr || !t_rowptr || !status || !workspace || (nnz > 0 && (!col || !t_col || !t_perm))) return ATTGNN_E_ARG;
    if (workspace_bytes < ws_bytes(n_cols, nnz)) return ATTGNN_E_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    TWs w = ws_carve(workspace, n_cols, nnz);
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n_cols + 1, THREADS)), dim3(THREADS), 0, s, w.counts, n_cols + 1,
                       w.status);
    if (n_rows > 0)
        hipLaunchKernelGGL(count_kernel, dim3(grid_for(n_rows, THREADS / 64)), dim3(THREADS), 0, s, rowptr, col,
                           n_rows, n_cols, nnz, w);
    size_t tb = w.temp_bytes;
    hipError_t e = rocprim::exclusive_scan(w.temp, tb, w.counts, t_rowptr, 0, size_t(n_cols + 1),
                                           rocprim::plus<int32_t>(), s);
    if (e != hipSuccess) return int(e);
    if (nnz > 0) {
        tb = w.temp_bytes;
        e = rocprim::radix_sort_pairs(w.temp, tb, col, w.keys_out, rocprim::counting_iterator<int32_t>(0), w.vals_out,
                                      size_t(nnz), 0, unsigned(bits_for(n_cols > 1 ? n_cols : 2)), s);
        if (e != hipSuccess) return int(e);
        hipLaunchKernelGGL(gather_kernel, dim3(grid_for(nnz, THREADS)), dim3(THREADS), 0, s, w, nnz, t_col, t_perm);
    }
    hipLaunchKernelGGL(status_kernel, dim3(1), dim3(64), 0, s, rowptr, n_rows, nnz, w, status);
    e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_map_remap_device(const int32_t *map, int64_t n, const int64_t *remap, int64_t remap_n,
                                       int32_t *out, attgnn_stream_t stream)
{
    if (n < 0 || remap_n < 0 || (n > 0 && (!map || !out || (remap_n > 0 && !remap)))) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    hipLaunchKernelGGL(remap_kernel, dim3(grid_for(n, THREADS)), dim3(THREADS), 0,
                       reinterpret_cast<hipStream_t>(stream), map, n, remap, remap_n, out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}
