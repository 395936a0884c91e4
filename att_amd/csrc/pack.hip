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
