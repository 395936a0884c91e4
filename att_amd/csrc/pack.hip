// Block-diagonal packing of many small graphs into one window-block structure
// (device side of the serving loop).
//
// No reference counterpart: the reference serves nothing and runs one graph
// per launch (end2end_ori/gcn_final/mgat_csr/test_mgat_csr.py:12-32).  Each
// cached graph holds its own arrays (rows padded to the window height so
// windows never straddle graphs); this kernel concatenates them into static,
// capacity-sized batch buffers with the offsets applied:
//   index arrays (column ids, window ids, CSR columns, edge ids): + offset,
//     -1 entries kept (padding slots);
//   running offsets (part offsets rp, CSR row pointers, int32; compact block
//     offsets, int64): + offset;
//   raw bytes (row masks, packed tile values, features): copied.
// The tail of every buffer beyond the batch is filled with "empty" entries
// (rp / row pointers repeat the total, window ids are -1, which marks a
// padding part the SpMM skips), so kernels launched with capacity sizes do no
// work there -- which lets one HIP graph per capacity bucket serve every batch.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>

#include "attgnn.h"

namespace {

// seg.kind: 0 = int32 + offset (keep -1), 1 = int32 + offset (all), 2 = bytes,
// 3 = int64 + offset
struct PackSeg {
    const void *src;
    void *dst;
    int64_t count;   // elements (int32 / int64) or bytes
    int64_t first;   // index of the segment's first chunk
    int32_t offset;
    int32_t kind;
};

// One workgroup per chunk of chunk_bytes: chunk c belongs to segment
// chunk_seg[c] (-1: nothing) and covers its items from (c - first) * per.
__global__ __launch_bounds__(256) void pack_kernel(const PackSeg *segs, int64_t nsegs, const int32_t *chunk_seg,
                                                   int64_t n_chunks, int64_t chunk_bytes)
{
    const int64_t c = blockIdx.x;
    if (c >= n_chunks) return;
    const int32_t si = chunk_seg[c];
    if (si < 0 || si >= nsegs) return;
    const PackSeg seg = segs[si];
    const int64_t unit = seg.kind == 2 ? 1 : seg.kind == 3 ? 8 : 4;
    const int64_t per = chunk_bytes / unit;
    const int64_t i0 = (c - seg.first) * per;
    if (i0 < 0 || i0 >= seg.count) return;
    const int64_t n = seg.count - i0 < per ? seg.count - i0 : per;
    if (seg.kind == 2) {
        const uint8_t *src = static_cast<const uint8_t *>(seg.src) + i0;
        uint8_t *dst = static_cast<uint8_t *>(seg.dst) + i0;
        const bool vec = (reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 16 == 0;
        const int64_t n16 = vec ? n / 16 : 0;
        const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
        uint4 *d4 = reinterpret_cast<uint4 *>(dst);
        // four independent 16-byte loads in flight per thread before the stores
        constexpr int U = 4;
        int64_t i = threadIdx.x;
        for (; i + (U - 1) * blockDim.x < n16; i += U * blockDim.x) {
            uint4 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = s4[i + u * blockDim.x];
#pragma unroll
            for (int u = 0; u < U; ++u) d4[i + u * blockDim.x] = v[u];
        }
        for (; i < n16; i += blockDim.x) d4[i] = s4[i];
        for (int64_t j = n16 * 16 + threadIdx.x; j < n; j += blockDim.x) dst[j] = src[j];
    } else if (seg.kind == 3) {
        const int64_t *src = seg.src ? static_cast<const int64_t *>(seg.src) + i0 : nullptr;
        int64_t *dst = static_cast<int64_t *>(seg.dst) + i0;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = (src ? src[i] : 0) + seg.offset;
    } else {
        const int32_t *src = seg.src ? static_cast<const int32_t *>(seg.src) + i0 : nullptr;
        int32_t *dst = static_cast<int32_t *>(seg.dst) + i0;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
            const int32_t v = src ? src[i] : 0;
            dst[i] = (seg.kind == 0 && v < 0) ? v : v + seg.offset;
        }
    }
}

}  // namespace

extern "C" int attgnn_pack_segments(const void *segs, int64_t nsegs, const int32_t *chunk_seg, int64_t n_chunks,
                                    int64_t chunk_bytes, attgnn_stream_t stream)
{
    if (nsegs < 0 || nsegs > INT32_MAX || n_chunks < 0 || n_chunks > INT32_MAX) return ATTGNN_E_ARG;
    if (chunk_bytes <= 0 || chunk_bytes % 16 != 0) return ATTGNN_E_ARG;
    if (n_chunks == 0 || nsegs == 0) return ATTGNN_OK;
    if (!segs || !chunk_seg) return ATTGNN_E_ARG;
    hipLaunchKernelGGL(pack_kernel, dim3(unsigned(n_chunks)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       static_cast<const PackSeg *>(segs), nsegs, chunk_seg, n_chunks, chunk_bytes);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int64_t attgnn_pack_segment_bytes(void) { return int64_t(sizeof(PackSeg)); }

// ---------------------------------------------------------------------------
// Host staging of one serving graph: the CSR padded to whole windows, the
// GCN normalising degrees and the features rounded to BF16, written into a
// pinned buffer in the device layout and uploaded with one copy.

namespace {

int64_t stage_align(int64_t v) { return (v + 255) / 256 * 256; }

// float bits -> bfloat16 bits, round to nearest even (as torch casts); NaN
// -> the quiet NaN 0x7FC0.  A branch-free loop over distinct arrays, which the
// host compiler vectorises.
void to_bf16(const uint32_t *__restrict__ src, uint16_t *__restrict__ dst, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t u = src[i];
        const uint32_t r = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
        dst[i] = uint16_t((u & 0x7fffffffu) > 0x7f800000u ? 0x7fc0u : r);
    }
}

}  // namespace

extern "C" int64_t attgnn_serve_stage_layout(int64_t n, int64_t nnz, int64_t feat, int window, int64_t *offsets)
{
    if (n < 0 || nnz < 0 || nnz > INT32_MAX || feat <= 0 || window <= 0 || !offsets) return ATTGNN_E_ARG;
    const int64_t n_pad = n ? (n + window - 1) / window * window : window;
    offsets[0] = 0;                                       // row pointers, int32 [n_pad + 1]
    offsets[1] = stage_align(4 * (n_pad + 1));            // column ids, int32 [nnz]
    offsets[2] = offsets[1] + stage_align(4 * nnz);       // degrees d_i, float [n_pad]
    offsets[3] = offsets[2] + stage_align(4 * n_pad);     // features, bf16 [n_pad][feat]
    offsets[4] = n_pad;
    return offsets[3] + stage_align(2 * n_pad * feat);
}

extern "C" int attgnn_serve_stage_host(const int64_t *indptr, const int32_t *indices, int64_t n, int64_t nnz,
                                       const float *features, int64_t feat, int window, void *out,
                                       int64_t out_bytes)
{
    int64_t off[5];
    const int64_t total = attgnn_serve_stage_layout(n, nnz, feat, window, off);
    if (total < 0) return int(total);
    if (!indptr || !out || out_bytes < total || (nnz > 0 && !indices)) return ATTGNN_E_ARG;
    if (indptr[0] != 0 || indptr[n] != nnz) return ATTGNN_E_ARG;
    const int64_t n_pad = off[4];
    uint8_t *p = static_cast<uint8_t *>(out);
    int32_t *rp = reinterpret_cast<int32_t *>(p + off[0]);
    float *d = reinterpret_cast<float *>(p + off[2]);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t a = indptr[i], b = indptr[i + 1];
        if (b < a) return ATTGNN_E_ARG;
        rp[i] = int32_t(a);
        // d_i = 1 / sqrt(row count) in float, 0 for an empty row
        d[i] = b > a ? 1.0f / std::sqrt(float(b - a)) : 0.0f;
    }
    for (int64_t i = n; i <= n_pad; ++i) rp[i] = int32_t(nnz);
    for (int64_t i = n; i < n_pad; ++i) d[i] = 0.0f;
    if (nnz > 0) std::memcpy(p + off[1], indices, size_t(nnz) * 4);
    uint16_t *x = reinterpret_cast<uint16_t *>(p + off[3]);
    const int64_t nx = n * feat;
    if (features) to_bf16(reinterpret_cast<const uint32_t *>(features), x, nx);
    std::memset(x + nx, 0, size_t((n_pad - n) * feat) * 2);
    return ATTGNN_OK;
}

extern "C" int attgnn_serve_stage(const int64_t *indptr, const int32_t *indices, int64_t n, int64_t nnz,
                                  const float *features, int64_t feat, int window, void *pinned,
                                  int64_t pinned_bytes, void *dev, attgnn_stream_t stream)
{
    if (!dev) return ATTGNN_E_ARG;
    const int rc = attgnn_serve_stage_host(indptr, indices, n, nnz, features, feat, window, pinned, pinned_bytes);
    if (rc != ATTGNN_OK) return rc;
    int64_t off[5];
    const int64_t total = attgnn_serve_stage_layout(n, nnz, feat, window, off);
    const hipError_t e = hipMemcpyAsync(dev, pinned, size_t(total), hipMemcpyHostToDevice,
                                        reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? ATTGNN_OK : int(e);
}
