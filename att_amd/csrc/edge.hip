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

// ---- all heads of a row in one 16-lane group --------------------------------
// The per-(row, head) arithmetic below is the one of the kernels above, step
// for step (same lane strides, same online update, same butterfly order), so
// the results are bitwise equal to theirs; the row's column ids are read once
// for all heads, and with node-major scores (element (h, i) at s[i*H + h]) the
// H terms of a neighbour are one vector load.

template <int H, bool NM>
__device__ __forceinline__ void load_heads(const float *s, int64_t hs, int64_t rs, int64_t i, float (&v)[H])
{
    if constexpr (NM && H == 8) {
        const f32x4 a = reinterpret_cast<const f32x4 *>(s + i * 8)[0];
        const f32x4 b = reinterpret_cast<const f32x4 *>(s + i * 8)[1];
#pragma unroll
        for (int h = 0; h < 4; ++h) { v[h] = a[h]; v[h + 4] = b[h]; }
    } else if constexpr (NM && H == 4) {
        const f32x4 a = *reinterpret_cast<const f32x4 *>(s + i * 4);
#pragma unroll
        for (int h = 0; h < 4; ++h) v[h] = a[h];
    } else if constexpr (NM && H == 2) {
        const u32x2 a = *reinterpret_cast<const u32x2 *>(s + i * 2);
        v[0] = __uint_as_float(a[0]); v[1] = __uint_as_float(a[1]);
    } else {
#pragma unroll
        for (int h = 0; h < H; ++h) v[h] = s[h * hs + i * rs];
    }
}

// Edges a lane handles per step (loads issued together): a hub row of ~1,500
// entries is ~90 steps of a 16-lane group, each a dependent col -> score load
// pair, so the loads of U steps are issued before their updates, which run in
// the same edge order as before.
template <int H>
constexpr int unroll_of() { return H >= 8 ? 4 : 8; }

template <int H, bool NM>
__global__ __launch_bounds__(256) void softmax_fwd_heads_kernel(SoftmaxArgs a)
{
    constexpr int U = unroll_of<H>();
    const int64_t i = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (i >= a.n) return;                      // whole 16-lane group exits together
    const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    float si[H], m[H], s[H];
    load_heads<H, NM>(a.s0, a.s_hs, a.s_rs, i, si);
#pragma unroll
    for (int h = 0; h < H; ++h) { m[h] = -INFINITY; s[h] = 0.f; }
    for (int64_t eb = e0 + l; eb < e1; eb += 16 * U) {
        float sj[U][H];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + 16 * u;
            load_heads<H, NM>(a.s1, a.s_hs, a.s_rs, e < e1 ? int64_t(a.col[e]) : 0, sj[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (eb + 16 * u >= e1) break;
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const float v = lrelu(si[h] + sj[u][h], a.slope);
                const float mn = fmaxf(m[h], v);
                s[h] = s[h] * __expf(m[h] - mn) + __expf(v - mn);
                m[h] = mn;
            }
        }
    }
    float lse[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const float M = group_max(m[h]);
        const float sh = (m[h] == -INFINITY) ? 0.f : s[h] * __expf(m[h] - M);
        lse[h] = M + __logf(group_sum(sh));
        if (l == 0) a.lse_out[h * a.n + i] = (e1 > e0) ? lse[h] : 0.f;
    }
    for (int64_t eb = e0 + l; eb < e1; eb += 16 * U) {
        float sj[U][H];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + 16 * u;
            load_heads<H, NM>(a.s1, a.s_hs, a.s_rs, e < e1 ? int64_t(a.col[e]) : 0, sj[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + 16 * u;
            if (e >= e1) break;
#pragma unroll
            for (int h = 0; h < H; ++h) {
                float p = __expf(lrelu(si[h] + sj[u][h], a.slope) - lse[h]);
                if (a.thresh) p = keep_edge(a.seed, h, e, a.thresh) ? p * a.keep_scale : 0.f;
                st(a.p, a.p_type, int64_t(h) * a.nnz + e, p);
            }
        }
    }
}

// EDGE_MAJOR: dz[e*H + h] (one vector store per edge) instead of dz[h*nnz + e].
template <int H, bool NM, bool EDGE_MAJOR>
__global__ __launch_bounds__(256) void softmax_bwd_heads_kernel(SoftmaxArgs a)
{
    constexpr int U = unroll_of<H>();
    const int64_t i = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (i >= a.n) return;
    const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    float si[H], lse[H], acc[H];
    load_heads<H, NM>(a.s0, a.s_hs, a.s_rs, i, si);
#pragma unroll
    for (int h = 0; h < H; ++h) { lse[h] = a.lse[h * a.n + i]; acc[h] = 0.f; }
    for (int64_t eb = e0 + l; eb < e1; eb += 16 * U) {
        float sj[U][H], g[U][H];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + 16 * u;
            const bool ok = e < e1;
            load_heads<H, NM>(a.s1, a.s_hs, a.s_rs, ok ? int64_t(a.col[e]) : 0, sj[u]);
#pragma unroll
            for (int h = 0; h < H; ++h) g[u][h] = ok ? ld(a.dp, a.dp_type, int64_t(h) * a.nnz + e) : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + 16 * u;
            if (e >= e1) break;
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const float p = __expf(lrelu(si[h] + sj[u][h], a.slope) - lse[h]);
                float gv = g[u][h];
                if (a.thresh) gv = keep_edge(a.seed, h, e, a.thresh) ? gv * a.keep_scale : 0.f;
                acc[h] += p * gv;
            }
        }
    }
    float S[H], ds[H];
#pragma unroll
    for (int h = 0; h < H; ++h) { S[h] = group_sum(acc[h]); ds[h] = 0.f; }
    for (int64_t eb = e0 + l; eb < e1; eb += 16 * U) {
        float sj[U][H], g[U][H];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + 16 * u;
            const bool ok = e < e1;
            load_heads<H, NM>(a.s1, a.s_hs, a.s_rs, ok ? int64_t(a.col[e]) : 0, sj[u]);
#pragma unroll
            for (int h = 0; h < H; ++h) g[u][h] = ok ? ld(a.dp, a.dp_type, int64_t(h) * a.nnz + e) : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + 16 * u;
            if (e >= e1) break;
            float dz[H];
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const float z = si[h] + sj[u][h];
                const float p = __expf(lrelu(z, a.slope) - lse[h]);
                float gv = g[u][h];
                if (a.thresh) gv = keep_edge(a.seed, h, e, a.thresh) ? gv * a.keep_scale : 0.f;
                dz[h] = p * (gv - S[h]) * (z > 0.f ? 1.f : a.slope);
                ds[h] += dz[h];
            }
            if constexpr (EDGE_MAJOR && H == 8) {
                reinterpret_cast<f32x4 *>(a.dz + e * 8)[0] = f32x4{dz[0], dz[1], dz[2], dz[3]};
                reinterpret_cast<f32x4 *>(a.dz + e * 8)[1] = f32x4{dz[4], dz[5], dz[6], dz[7]};
            } else if constexpr (EDGE_MAJOR && H == 4) {
                *reinterpret_cast<f32x4 *>(a.dz + e * 4) = f32x4{dz[0], dz[1], dz[2], dz[3]};
            } else if constexpr (EDGE_MAJOR && H == 2) {
                *reinterpret_cast<u32x2 *>(a.dz + e * 2) = u32x2{__float_as_uint(dz[0]), __float_as_uint(dz[1])};
            } else if constexpr (EDGE_MAJOR) {
                a.dz[e] = dz[0];
            } else {
#pragma unroll
                for (int h = 0; h < H; ++h) a.dz[int64_t(h) * a.nnz + e] = dz[h];
            }
        }
    }
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const float v = group_sum(ds[h]);
        if (l == 0) a.ds0[h * a.n + i] = v;
    }
}

// out[h*n + j] = sum over row j of vals(h, perm[e]) (perm -1 or out of range =
// skip), vals(h, q) at vals[q*H + h] (EDGE_MAJOR) or vals[h*nnz + q]: the
// column sums of dz through the transpose permutation, with the lane order of
// gather_rowsum_kernel and the permutation read once for all heads.
template <int H, bool EDGE_MAJOR>
__global__ __launch_bounds__(256) void gather_rowsum_heads_kernel(const int32_t *rowptr, const int64_t *perm,
                                                                  int64_t n, int64_t nnz, const float *vals,
                                                                  float *out)
{
    constexpr int U = unroll_of<H>();
    const int64_t j = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (j >= n) return;
    float s[H];
#pragma unroll
    for (int h = 0; h < H; ++h) s[h] = 0.f;
    const int64_t b = rowptr[j], e1 = rowptr[j + 1];
    for (int64_t eb = b + l; eb < e1; eb += 16 * U) {
        int64_t q[U];
        float v[U][H];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + 16 * u;
            q[u] = e < e1 ? perm[e] : -1;
            const bool ok = q[u] >= 0 && q[u] < nnz;
            load_heads<H, EDGE_MAJOR>(vals, nnz, 1, ok ? q[u] : 0, v[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (q[u] >= 0 && q[u] < nnz) {
#pragma unroll
                for (int h = 0; h < H; ++h) s[h] += v[u][h];
            }
        }
    }
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const float v = group_sum(s[h]);
        if (l == 0) out[h * n + j] = v;
    }
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
        if (q >= 0 && q < nnz) s += vals[int64_t(h) * nnz + q];
    }
    s = group_sum(s);
    if (l == 0) out[h * n + j] = s;
}

bool float_type(int t) { return t == ATTGNN_F32 || t == ATTGNN_F16 || t == ATTGNN_BF16; }

uint32_t drop_threshold(float q) { return q <= 0.f ? 0u : uint32_t(std::min(double(q), 0.999999) * 4294967296.0); }

// Heads handled by the all-heads kernels; others use the per-(row, head) ones.
bool heads_kernel_ok(int64_t heads) { return heads == 1 || heads == 2 || heads == 4 || heads == 8; }

// Node-major scores, aligned for the vector loads of load_heads.
bool node_major(int64_t heads, int64_t hs, int64_t rs, const float *s0, const float *s1)
{
    if (heads == 1) return false;
    const uintptr_t al = uintptr_t(heads >= 4 ? 16 : 8) - 1;
    return hs == 1 && rs == heads && (reinterpret_cast<uintptr_t>(s0) & al) == 0 &&
           (reinterpret_cast<uintptr_t>(s1) & al) == 0;
}

hipError_t launch_fwd_heads(const SoftmaxArgs &a, bool nm, hipStream_t s)
{
    const dim3 grid(unsigned((a.n * 16 + 255) / 256)), block(256);
    switch (a.heads * 2 + int(nm)) {
    case 2: hipLaunchKernelGGL((softmax_fwd_heads_kernel<1, false>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((softmax_fwd_heads_kernel<2, false>), grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL((softmax_fwd_heads_kernel<2, true>), grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL((softmax_fwd_heads_kernel<4, false>), grid, block, 0, s, a); break;
    case 9: hipLaunchKernelGGL((softmax_fwd_heads_kernel<4, true>), grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL((softmax_fwd_heads_kernel<8, false>), grid, block, 0, s, a); break;
    case 17: hipLaunchKernelGGL((softmax_fwd_heads_kernel<8, true>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool EM>
hipError_t launch_bwd_heads(const SoftmaxArgs &a, bool nm, hipStream_t s)
{
    const dim3 grid(unsigned((a.n * 16 + 255) / 256)), block(256);
    switch (a.heads * 2 + int(nm)) {
    case 2: hipLaunchKernelGGL((softmax_bwd_heads_kernel<1, false, EM>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((softmax_bwd_heads_kernel<2, false, EM>), grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL((softmax_bwd_heads_kernel<2, true, EM>), grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL((softmax_bwd_heads_kernel<4, false, EM>), grid, block, 0, s, a); break;
    case 9: hipLaunchKernelGGL((softmax_bwd_heads_kernel<4, true, EM>), grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL((softmax_bwd_heads_kernel<8, false, EM>), grid, block, 0, s, a); break;
    case 17: hipLaunchKernelGGL((softmax_bwd_heads_kernel<8, true, EM>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool EM>
hipError_t launch_gather_heads(const int32_t *rowptr, const int64_t *perm, int64_t n, int64_t nnz, int heads,
                               const float *vals, float *out, hipStream_t s)
{
    const dim3 grid(unsigned((n * 16 + 255) / 256)), block(256);
    switch (heads) {
    case 1: hipLaunchKernelGGL((gather_rowsum_heads_kernel<1, EM>), grid, block, 0, s, rowptr, perm, n, nnz, vals, out); break;
    case 2: hipLaunchKernelGGL((gather_rowsum_heads_kernel<2, EM>), grid, block, 0, s, rowptr, perm, n, nnz, vals, out); break;
    case 4: hipLaunchKernelGGL((gather_rowsum_heads_kernel<4, EM>), grid, block, 0, s, rowptr, perm, n, nnz, vals, out); break;
    case 8: hipLaunchKernelGGL((gather_rowsum_heads_kernel<8, EM>), grid, block, 0, s, rowptr, perm, n, nnz, vals, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

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
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (heads_kernel_ok(heads)) {
        const hipError_t e = launch_fwd_heads(a, node_major(heads, s_head_stride, s_row_stride, s0, s1), s);
        return e == hipSuccess ? ATTGNN_OK : int(e);
    }
    const int64_t threads = n * heads * 16;
    hipLaunchKernelGGL(edge_softmax_fwd_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

namespace {

SoftmaxArgs bwd_args(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz, int64_t heads,
                     const float *s0, const float *s1, int64_t s_head_stride, int64_t s_row_stride, const float *lse,
                     const void *dp, int dp_type, float slope, float drop_p, uint64_t seed, float *dz, float *ds0)
{
    SoftmaxArgs a{};
    a.rowptr = rowptr; a.col = col; a.s0 = s0; a.s1 = s1; a.lse = lse; a.dp = static_cast<const uint8_t *>(dp);
    a.dz = dz; a.ds0 = ds0; a.n = n; a.nnz = nnz; a.s_hs = s_head_stride; a.s_rs = s_row_stride;
    a.heads = int(heads); a.dp_type = dp_type; a.slope = slope; a.thresh = drop_threshold(drop_p);
    a.keep_scale = 1.f / (1.f - drop_p); a.seed = seed;
    return a;
}

}  // namespace

extern "C" int attgnn_edge_softmax_bwd(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz,
                                       int64_t heads, const float *s0, const float *s1, int64_t s_head_stride,
                                       int64_t s_row_stride, const float *lse, const void *dp, int dp_type,
                                       float slope, float drop_p, uint64_t seed, float *dz, float *ds0,
                                       attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0 || !float_type(dp_type) || drop_p < 0.f || drop_p >= 1.f) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !s0 || !s1 || !lse || !ds0 || (nnz > 0 && (!col || !dp || !dz))) return ATTGNN_E_ARG;
    const SoftmaxArgs a = bwd_args(rowptr, col, n, nnz, heads, s0, s1, s_head_stride, s_row_stride, lse, dp, dp_type,
                                   slope, drop_p, seed, dz, ds0);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (heads_kernel_ok(heads)) {
        const hipError_t e = launch_bwd_heads<false>(a, node_major(heads, s_head_stride, s_row_stride, s0, s1), s);
        return e == hipSuccess ? ATTGNN_OK : int(e);
    }
    const int64_t threads = n * heads * 16;
    hipLaunchKernelGGL(edge_softmax_bwd_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int64_t attgnn_edge_softmax_bwd_sums_workspace_elems(int64_t nnz, int64_t heads)
{
    if (nnz < 0 || heads <= 0) return -1;
    return nnz * heads;
}

extern "C" int attgnn_edge_softmax_bwd_sums(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz,
                                            int64_t heads, const float *s0, const float *s1,
                                            int64_t s_head_stride, int64_t s_row_stride, const float *lse,
                                            const void *dp, int dp_type, float slope, float drop_p, uint64_t seed,
                                            const int32_t *t_rowptr, const int64_t *t_perm, float *ds0, float *ds1,
                                            float *workspace, int64_t workspace_elems, attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0 || !float_type(dp_type) || drop_p < 0.f || drop_p >= 1.f) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !t_rowptr || !s0 || !s1 || !lse || !ds0 || !ds1) return ATTGNN_E_ARG;
    if (nnz > 0 && (!col || !dp || !t_perm || !workspace)) return ATTGNN_E_ARG;
    if (workspace_elems < nnz * heads) return ATTGNN_E_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const bool em = heads_kernel_ok(heads) && (reinterpret_cast<uintptr_t>(workspace) & 15) == 0;
    const SoftmaxArgs a = bwd_args(rowptr, col, n, nnz, heads, s0, s1, s_head_stride, s_row_stride, lse, dp, dp_type,
                                   slope, drop_p, seed, workspace, ds0);
    hipError_t e;
    if (em) {
        // dz edge-major in the workspace, then column sums reading one record per edge
        e = launch_bwd_heads<true>(a, node_major(heads, s_head_stride, s_row_stride, s0, s1), s);
        if (e == hipSuccess) e = launch_gather_heads<true>(t_rowptr, t_perm, n, nnz, int(heads), workspace, ds1, s);
    } else {
        const int64_t threads = n * heads * 16;
        hipLaunchKernelGGL(edge_softmax_bwd_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s, a);
        hipLaunchKernelGGL(gather_rowsum_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s, t_rowptr,
                           t_perm, n, nnz, int(heads), static_cast<const float *>(workspace), ds1);
        e = hipGetLastError();
    }
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_gather_rowsum(const int32_t *rowptr, const int64_t *perm, int64_t n, int64_t nnz,
                                    int64_t heads, const float *vals, float *out, attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !out || (nnz > 0 && (!perm || !vals))) return ATTGNN_E_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (heads_kernel_ok(heads)) {
        const hipError_t e = launch_gather_heads<false>(rowptr, perm, n, nnz, int(heads), vals, out, s);
        return e == hipSuccess ? ATTGNN_OK : int(e);
    }
    const int64_t threads = n * heads * 16;
    hipLaunchKernelGGL(gather_rowsum_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s, rowptr, perm,
                       n, nnz, int(heads), vals, out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}
