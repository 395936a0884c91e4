// Fused per-edge kernels of the attention layers (gfx950), over CSR rows.
//
// Reference counterpart: the chain of separate torch ops and extension calls in
// end2end_head/gat_final/mgat_csr/mgat_conv.py:180-197 (LeakyReLU, exp, rowsum
// through ATT_SpMM_v3.spmm_fp16_m_h_rowsum, dropout, divide) and its backward
// (:34-69, :91-151).  Here:
//   forward : logits z = s0[h,i] + s1[h,j]  ->  LeakyReLU  ->  row softmax (max
//             subtracted, FP32)  ->  dropout  ->  p[h,e], plus lse[h,i];
//   backward: recompute p from (s0, s1, lse) instead of storing it, apply the
//             dropout mask again from its counter-based hash, softmax and
//             LeakyReLU backward, and the row sums ds0[h,i] = sum_j dz[h,(i,j)].
// One 16-lane group per (row, head); a row of any length is walked in strides
// of 16 with an online (max, sum) pair, so a hub row costs one group.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "attgnn.h"
#include "mfma.hpp"

using namespace attgnn;

namespace {

__device__ __forceinline__ float lrelu(float z, float slope) { return z > 0.f ? z : z * slope; }

// Counter-based keep decision for dropout: deterministic in (seed, head, edge),
// so the backward pass regenerates the forward mask without storing it.
__device__ __forceinline__ bool keep_edge(uint64_t seed, int64_t h, int64_t e, uint32_t thresh)
{
    uint64_t x = seed ^ (uint64_t(h) * 0x9E3779B97F4A7C15ull) ^ (uint64_t(e) * 0xC2B2AE3D27D4EB4Full);
    x ^= x >> 31; x *= 0x7FB5D329728EA185ull;
    x ^= x >> 27; x *= 0x81DADEF4BC2DD44Dull;
    x ^= x >> 33;
    return uint32_t(x) >= thresh;
}

__device__ __forceinline__ float ld(const uint8_t *p, int type, int64_t i)
{
    switch (type) {
    case ATTGNN_F32: return reinterpret_cast<const float *>(p)[i];
    case ATTGNN_F16: return float(reinterpret_cast<const _Float16 *>(p)[i]);
    default: return __uint_as_float(uint32_t(reinterpret_cast<const uint16_t *>(p)[i]) << 16);
    }
}

__device__ __forceinline__ void st(uint8_t *p, int type, int64_t i, float v)
{
    switch (type) {
    case ATTGNN_F32: reinterpret_cast<float *>(p)[i] = v; break;
    case ATTGNN_F16: reinterpret_cast<_Float16 *>(p)[i] = _Float16(v); break;
    default: reinterpret_cast<uint16_t *>(p)[i] = uint16_t(bf16_bits_rne(v)); break;
    }
}

__device__ __forceinline__ float group_max(float v)
{
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 16));
    return v;
}
__device__ __forceinline__ float group_sum(float v)
{
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) v += __shfl_xor(v, o, 16);
    return v;
}

struct SoftmaxArgs {
    const int32_t *rowptr, *col;
    const float *s0, *s1;      // [heads][n] node terms of the logits
    const float *lse;          // backward: [heads][n]
    const uint8_t *dp;         // backward: dL/dp (after dropout), [heads][nnz] of dp_type
    uint8_t *p;                // forward output [heads][nnz] of p_type
    float *lse_out;            // forward output
    float *dz;                 // backward output [heads][nnz]
    float *ds0;                // backward output [heads][n]
    int64_t n, nnz, s_hs, s_rs;
    int heads, p_type, dp_type;
    float slope, keep_scale;   // keep_scale = 1/(1-q)
    uint32_t thresh;           // q * 2^32
    uint64_t seed;
};

__global__ __launch_bounds__(256) void edge_softmax_fwd_kernel(SoftmaxArgs a)
{
    const int64_t gid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (gid >= a.n * a.heads) return;          // whole 16-lane group exits together
    const int64_t i = gid / a.heads;
    const int h = int(gid - i * a.heads);
    const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    const float si = a.s0[h * a.s_hs + i * a.s_rs];
    float m = -INFINITY, s = 0.f;
    for (int64_t e = e0 + l; e < e1; e += 16) {
        const float v = lrelu(si + a.s1[h * a.s_hs + int64_t(a.col[e]) * a.s_rs], a.slope);
        const float mn = fmaxf(m, v);
        s = s * __expf(m - mn) + __expf(v - mn);
        m = mn;
    }
    const float M = group_max(m);
    s = (m == -INFINITY) ? 0.f : s * __expf(m - M);
    const float S = group_sum(s);
    const float lse = M + __logf(S);
    if (l == 0) a.lse_out[h * a.n + i] = (e1 > e0) ? lse : 0.f;
    for (int64_t e = e0 + l; e < e1; e += 16) {
        const float v = lrelu(si + a.s1[h * a.s_hs + int64_t(a.col[e]) * a.s_rs], a.slope);
        float p = __expf(v - lse);
        if (a.thresh) p = keep_edge(a.seed, h, e, a.thresh) ? p * a.keep_scale : 0.f;
        st(a.p, a.p_type, int64_t(h) * a.nnz + e, p);
    }
}

__global__ __launch_bounds__(256) void edge_softmax_bwd_kernel(SoftmaxArgs a)
{
    const int64_t gid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (gid >= a.n * a.heads) return;
    const int64_t i = gid / a.heads;
    const int h = int(gid - i * a.heads);
    const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    const float si = a.s0[h * a.s_hs + i * a.s_rs];
    const float lse = a.lse[h * a.n + i];
    float acc = 0.f;  // sum_k p_k g_k, g = dL/dp before dropout
    for (int64_t e = e0 + l; e < e1; e += 16) {
        const float p = __expf(lrelu(si + a.s1[h * a.s_hs + int64_t(a.col[e]) * a.s_rs], a.slope) - lse);
        float g = ld(a.dp, a.dp_type, int64_t(h) * a.nnz + e);
        if (a.thresh) g = keep_edge(a.seed, h, e, a.thresh) ? g * a.keep_scale : 0.f;
        acc += p * g;
    }
    const float S = group_sum(acc);
    float ds = 0.f;
    for (int64_t e = e0 + l; e < e1; e += 16) {
        const float z = si + a.s1[h * a.s_hs + int64_t(a.col[e]) * a.s_rs];
        const float p = __expf(lrelu(z, a.slope) - lse);
        float g = ld(a.dp, a.dp_type, int64_t(h) * a.nnz + e);
        if (a.thresh) g = keep_edge(a.seed, h, e, a.thresh) ? g * a.keep_scale : 0.f;
        const float dz = p * (g - S) * (z > 0.f ? 1.f : a.slope);
        a.dz[int64_t(h) * a.nnz + e] = dz;
        ds += dz;
    }
    ds = group_sum(ds);
    if (l == 0) a.ds0[h * a.n + i] = ds;
}

// out[h*n + j] = sum over row j of vals[h*nnz + perm[e]] (perm -1 = skip).
__global__ __launch_bounds__(256) void gather_rowsum_kernel(const int32_t *rowptr, const int64_t *perm, int64_t n,
                                                            int64_t nnz, int heads, const float *vals, float *out)
{
    const int64_t gid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (gid >= n * heads) return;
    const int64_t j = gid / heads;
    const int h = int(gid - j * heads);
    float s = 0.f;
    for (int64_t e = rowptr[j] + l; e < rowptr[j + 1]; e += 16) {
        const int64_t q = perm[e];
        if (q >= 0) s += vals[int64_t(h) * nnz + q];
    }
    s = group_sum(s);
    if (l == 0) out[h * n + j] = s;
}

bool float_type(int t) { return t == ATTGNN_F32 || t == ATTGNN_F16 || t == ATTGNN_BF16; }

uint32_t drop_threshold(float q) { return q <= 0.f ? 0u : uint32_t(std::min(double(q), 0.999999) * 4294967296.0); }

}  // namespace

extern "C" int attgnn_edge_softmax_fwd(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz,
                                       int64_t heads, const float *s0, const float *s1, int64_t s_head_stride,
                                       int64_t s_row_stride, float slope, float drop_p, uint64_t seed, void *p,
                                       int p_type, float *lse, attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0 || !float_type(p_type) || drop_p < 0.f || drop_p >= 1.f) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !s0 || !s1 || !lse || (nnz > 0 && (!col || !p))) return ATTGNN_E_ARG;
    SoftmaxArgs a{};
    a.rowptr = rowptr; a.col = col; a.s0 = s0; a.s1 = s1; a.p = static_cast<uint8_t *>(p); a.lse_out = lse;
    a.n = n; a.nnz = nnz; a.s_hs = s_head_stride; a.s_rs = s_row_stride; a.heads = int(heads); a.p_type = p_type;
    a.slope = slope; a.thresh = drop_threshold(drop_p); a.keep_scale = 1.f / (1.f - drop_p); a.seed = seed;
    const int64_t threads = n * heads * 16;
    hipLaunchKernelGGL(edge_softmax_fwd_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_edge_softmax_bwd(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz,
                                       int64_t heads, const float *s0, const float *s1, int64_t s_head_stride,
                                       int64_t s_row_stride, const float *lse, const void *dp, int dp_type,
                                       float slope, float drop_p, uint64_t seed, float *dz, float *ds0,
                                       attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0 || !float_type(dp_type) || drop_p < 0.f || drop_p >= 1.f) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !s0 || !s1 || !lse || !ds0 || (nnz > 0 && (!col || !dp || !dz))) return ATTGNN_E_ARG;
    SoftmaxArgs a{};
    a.rowptr = rowptr; a.col = col; a.s0 = s0; a.s1 = s1; a.lse = lse; a.dp = static_cast<const uint8_t *>(dp);
    a.dz = dz; a.ds0 = ds0; a.n = n; a.nnz = nnz; a.s_hs = s_head_stride; a.s_rs = s_row_stride;
    a.heads = int(heads); a.dp_type = dp_type; a.slope = slope; a.thresh = drop_threshold(drop_p);
    a.keep_scale = 1.f / (1.f - drop_p); a.seed = seed;
    const int64_t threads = n * heads * 16;
    hipLaunchKernelGGL(edge_softmax_bwd_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_gather_rowsum(const int32_t *rowptr, const int64_t *perm, int64_t n, int64_t nnz,
                                    int64_t heads, const float *vals, float *out, attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !out || (nnz > 0 && (!perm || !vals))) return ATTGNN_E_ARG;
    const int64_t threads = n * heads * 16;
    hipLaunchKernelGGL(gather_rowsum_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), rowptr, perm, n, nnz, int(heads), vals, out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}
