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
#include <climits>
#include <cstdint>

#include "attgnn.h"
#include "attn.hpp"
#include "mfma.hpp"

using namespace attgnn;

namespace {

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
    int64_t long_min;          // rows with more entries are left to the workgroup kernels
    const int32_t *long_rows;  // those rows (n_long of them)
    int64_t n_long;
    int heads, p_type, dp_type;
    int dz_edge_major;         // backward: dz[e*heads + h] instead of dz[h*nnz + e]
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
    if (e1 - e0 > a.long_min) return;          // a workgroup of its own takes it
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
    if (!a.p) return;                          // log-sum-exp only
    for (int64_t e = e0 + l; e < e1; e += 16) {
        const float p = attn_prob(si + a.s1[h * a.s_hs + int64_t(a.col[e]) * a.s_rs], lse, a.slope, a.thresh,
                                  a.keep_scale, a.seed, h, e);
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
    if (e1 - e0 > a.long_min) return;
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
// ---- rows with many entries: one 256-thread workgroup per (row, head) -------
// A 16-lane group walks a row serially, so a hub row of ~1,500 entries (power-
// law graphs) outlasts every other row by far; rows above long_min entries are
// skipped by the group kernels and taken by workgroups: workgroup b of the
// long part handles row long_rows[b / heads], head b % heads.  Each thread
// folds a 256-strided slice of the row; the block reduction is a butterfly
// inside each wave, then the four wave results in a fixed order, so the result
// is deterministic.  The all-heads kernels run these workgroups first in the
// same launch (so the hub rows overlap the short ones); the per-head kernels
// launch them separately.  Either way the same function computes them.

constexpr int LONG_THREADS = 256;

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// MAX: block maximum, else block sum; every thread gets the result.
template <bool MAX>
__device__ __forceinline__ float block_reduce(float v, float *red)
{
    v = MAX ? wave_max(v) : wave_sum(v);
    __syncthreads();                      // red[] free from the previous reduction
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = red[0];
#pragma unroll
    for (int w = 1; w < LONG_THREADS / 64; ++w) r = MAX ? fmaxf(r, red[w]) : r + red[w];
    return r;
}

// (row, head) of long-part workgroup b; false when the row is not long (then
// the group kernel has written it).  Uniform over the workgroup.
__device__ __forceinline__ bool long_task(const SoftmaxArgs &a, int64_t b, int64_t &i, int &h, int64_t &e0,
                                          int64_t &e1)
{
    i = a.long_rows[b / a.heads];
    h = int(b % a.heads);
    if (i < 0 || i >= a.n) return false;
    e0 = a.rowptr[i];
    e1 = a.rowptr[i + 1];
    return e1 - e0 > a.long_min;
}

__device__ __forceinline__ void softmax_fwd_long(const SoftmaxArgs &a, int64_t b, float *red)
{
    int64_t i, e0 = 0, e1 = 0;
    int h;
    if (!long_task(a, b, i, h, e0, e1)) return;
    const int t = threadIdx.x;
    const float si = a.s0[h * a.s_hs + i * a.s_rs];
    float m = -INFINITY, s = 0.f;
    for (int64_t e = e0 + t; e < e1; e += LONG_THREADS) {
        const float v = lrelu(si + a.s1[h * a.s_hs + int64_t(a.col[e]) * a.s_rs], a.slope);
        const float mn = fmaxf(m, v);
        s = s * __expf(m - mn) + __expf(v - mn);
        m = mn;
    }
    const float M = block_reduce<true>(m, red);
    const float S = block_reduce<false>((m == -INFINITY) ? 0.f : s * __expf(m - M), red);
    const float lse = M + __logf(S);
    if (t == 0) a.lse_out[h * a.n + i] = lse;
    if (!a.p) return;
    for (int64_t e = e0 + t; e < e1; e += LONG_THREADS) {
        const float p = attn_prob(si + a.s1[h * a.s_hs + int64_t(a.col[e]) * a.s_rs], lse, a.slope, a.thresh,
                                  a.keep_scale, a.seed, h, e);
        st(a.p, a.p_type, int64_t(h) * a.nnz + e, p);
    }
}

__device__ __forceinline__ void softmax_bwd_long(const SoftmaxArgs &a, int64_t b, float *red)
{
    int64_t i, e0 = 0, e1 = 0;
    int h;
    if (!long_task(a, b, i, h, e0, e1)) return;
    const int t = threadIdx.x;
    const float si = a.s0[h * a.s_hs + i * a.s_rs];
    const float lse = a.lse[h * a.n + i];
    float acc = 0.f;
    for (int64_t e = e0 + t; e < e1; e += LONG_THREADS) {
        const float p = __expf(lrelu(si + a.s1[h * a.s_hs + int64_t(a.col[e]) * a.s_rs], a.slope) - lse);
        float g = ld(a.dp, a.dp_type, int64_t(h) * a.nnz + e);
        if (a.thresh) g = keep_edge(a.seed, h, e, a.thresh) ? g * a.keep_scale : 0.f;
        acc += p * g;
    }
    const float S = block_reduce<false>(acc, red);
    float ds = 0.f;
    for (int64_t e = e0 + t; e < e1; e += LONG_THREADS) {
        const float z = si + a.s1[h * a.s_hs + int64_t(a.col[e]) * a.s_rs];
        const float p = __expf(lrelu(z, a.slope) - lse);
        float g = ld(a.dp, a.dp_type, int64_t(h) * a.nnz + e);
        if (a.thresh) g = keep_edge(a.seed, h, e, a.thresh) ? g * a.keep_scale : 0.f;
        const float dz = p * (g - S) * (z > 0.f ? 1.f : a.slope);
        a.dz[a.dz_edge_major ? e * a.heads + h : int64_t(h) * a.nnz + e] = dz;
        ds += dz;
    }
    ds = block_reduce<false>(ds, red);
    if (t == 0) a.ds0[h * a.n + i] = ds;
}

__global__ __launch_bounds__(LONG_THREADS) void softmax_fwd_long_kernel(SoftmaxArgs a)
{
    __shared__ float red[LONG_THREADS / 64];
    softmax_fwd_long(a, blockIdx.x, red);
}

__global__ __launch_bounds__(LONG_THREADS) void softmax_bwd_long_kernel(SoftmaxArgs a)
{
    __shared__ float red[LONG_THREADS / 64];
    softmax_bwd_long(a, blockIdx.x, red);
}

// Column sums of the listed long rows: out[h*n + j] with vals(h, q) at
// vals[q*heads + h] (edge_major) or vals[h*nnz + q].
struct GatherArgs {
    const int32_t *rowptr;
    const int64_t *perm;
    const float *vals;
    float *out;
    const int32_t *long_rows;
    int64_t n, nnz, long_min, n_long;
    int heads, edge_major;
};

__device__ __forceinline__ void gather_long(const GatherArgs &g, int64_t b, float *red)
{
    const int64_t j = g.long_rows[b / g.heads];
    const int h = int(b % g.heads);
    if (j < 0 || j >= g.n) return;
    const int64_t e0 = g.rowptr[j], e1 = g.rowptr[j + 1];
    if (e1 - e0 <= g.long_min) return;
    float s = 0.f;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += LONG_THREADS) {
        const int64_t q = g.perm[e];
        if (q >= 0 && q < g.nnz) s += g.vals[g.edge_major ? q * g.heads + h : int64_t(h) * g.nnz + q];
    }
    s = block_reduce<false>(s, red);
    if (threadIdx.x == 0) g.out[h * g.n + j] = s;
}

__global__ __launch_bounds__(LONG_THREADS) void gather_rowsum_long_kernel(GatherArgs g)
{
    __shared__ float red[LONG_THREADS / 64];
    gather_long(g, blockIdx.x, red);
}

// ---- all heads of a row in one 16-lane group (short rows) -------------------
// Edges a lane handles per step (loads issued together): the loads of U steps
// go out before their updates, which run in the same edge order as before.  A
// row of at most 16*U entries takes one step, and its second pass reuses the
// values loaded by the first instead of loading them again.
template <int H>
constexpr int unroll_of() { return H >= 8 ? 4 : 8; }

template <int H, bool NM>
__global__ __launch_bounds__(256) void softmax_fwd_heads_kernel(SoftmaxArgs a)
{
    constexpr int U = unroll_of<H>();
    __shared__ float red[LONG_THREADS / 64];
    const int64_t lb = a.n_long * a.heads;
    if (int64_t(blockIdx.x) < lb) {            // long rows first, a workgroup per (row, head)
        softmax_fwd_long(a, blockIdx.x, red);
        return;
    }
    const int64_t i = ((int64_t(blockIdx.x) - lb) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (i >= a.n) return;                      // whole 16-lane group exits together
    const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    if (e1 - e0 > a.long_min) return;          // a workgroup per head takes it
    const bool one_step = e1 - e0 <= 16 * U;
    float si[H], m[H], s[H];
    load_heads<H, NM>(a.s0, a.s_hs, a.s_rs, i, si);
#pragma unroll
    for (int h = 0; h < H; ++h) { m[h] = -INFINITY; s[h] = 0.f; }
    float sj[U][H];
    for (int64_t eb = e0 + l; eb < e1; eb += 16 * U) {
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
    if (!a.p) return;                          // log-sum-exp only
    for (int64_t eb = e0 + l; eb < e1; eb += 16 * U) {
        if (!one_step) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t e = eb + 16 * u;
                load_heads<H, NM>(a.s1, a.s_hs, a.s_rs, e < e1 ? int64_t(a.col[e]) : 0, sj[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + 16 * u;
            if (e >= e1) break;
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const float p = attn_prob(si[h] + sj[u][h], lse[h], a.slope, a.thresh, a.keep_scale, a.seed, h, e);
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
    __shared__ float red[LONG_THREADS / 64];
    const int64_t lb = a.n_long * a.heads;
    if (int64_t(blockIdx.x) < lb) {
        softmax_bwd_long(a, blockIdx.x, red);
        return;
    }
    const int64_t i = ((int64_t(blockIdx.x) - lb) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (i >= a.n) return;
    const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    if (e1 - e0 > a.long_min) return;
    const bool one_step = e1 - e0 <= 16 * U;
    float si[H], lse[H], acc[H];
    load_heads<H, NM>(a.s0, a.s_hs, a.s_rs, i, si);
#pragma unroll
    for (int h = 0; h < H; ++h) { lse[h] = a.lse[h * a.n + i]; acc[h] = 0.f; }
    float sj[U][H], g[U][H];
    for (int64_t eb = e0 + l; eb < e1; eb += 16 * U) {
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
        if (!one_step) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t e = eb + 16 * u;
                const bool ok = e < e1;
                load_heads<H, NM>(a.s1, a.s_hs, a.s_rs, ok ? int64_t(a.col[e]) : 0, sj[u]);
#pragma unroll
                for (int h = 0; h < H; ++h) g[u][h] = ok ? ld(a.dp, a.dp_type, int64_t(h) * a.nnz + e) : 0.f;
            }
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
// gather_rowsum_kernel and the permutation read once for all heads.  The first
// n_long*H workgroups take the long rows.
template <int H, bool EDGE_MAJOR>
__global__ __launch_bounds__(256) void gather_rowsum_heads_kernel(GatherArgs ga)
{
    constexpr int U = unroll_of<H>();
    __shared__ float red[LONG_THREADS / 64];
    const int64_t lb = ga.n_long * H;
    if (int64_t(blockIdx.x) < lb) {
        gather_long(ga, blockIdx.x, red);
        return;
    }
    const int64_t j = ((int64_t(blockIdx.x) - lb) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (j >= ga.n) return;
    float s[H];
#pragma unroll
    for (int h = 0; h < H; ++h) s[h] = 0.f;
    const int64_t b = ga.rowptr[j], e1 = ga.rowptr[j + 1];
    if (e1 - b > ga.long_min) return;
    for (int64_t eb = b + l; eb < e1; eb += 16 * U) {
        int64_t q[U];
        float v[U][H];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = eb + 16 * u;
            q[u] = e < e1 ? ga.perm[e] : -1;
            const bool ok = q[u] >= 0 && q[u] < ga.nnz;
            load_heads<H, EDGE_MAJOR>(ga.vals, ga.nnz, 1, ok ? q[u] : 0, v[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (q[u] >= 0 && q[u] < ga.nnz) {
#pragma unroll
                for (int h = 0; h < H; ++h) s[h] += v[u][h];
            }
        }
    }
#pragma unroll
    for (int h = 0; h < H; ++h) {
        const float v = group_sum(s[h]);
        if (l == 0) ga.out[h * ga.n + j] = v;
    }
}

// out[h*n + j] = sum over row j of vals[h*nnz + perm[e]] (perm -1 = skip).
__global__ __launch_bounds__(256) void gather_rowsum_kernel(const int32_t *rowptr, const int64_t *perm, int64_t n,
                                                            int64_t nnz, int heads, const float *vals, float *out,
                                                            int64_t long_min)
{
    const int64_t gid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (gid >= n * heads) return;
    const int64_t j = gid / heads;
    const int h = int(gid - j * heads);
    if (int64_t(rowptr[j + 1]) - rowptr[j] > long_min) return;
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

// Grid of the all-heads kernels: the long-row workgroups, then 16 rows per workgroup.
dim3 heads_grid(const SoftmaxArgs &a) { return dim3(unsigned(a.n_long * a.heads + (a.n * 16 + 255) / 256)); }

hipError_t launch_fwd_heads(const SoftmaxArgs &a, bool nm, hipStream_t s)
{
    const dim3 grid = heads_grid(a), block(256);
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
    const dim3 grid = heads_grid(a), block(256);
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

// Column sums: the all-heads kernel (1, 2, 4 or 8 heads; long rows in the same
// launch) or the per-head kernel followed by a workgroup per (long row, head).
hipError_t launch_gather(const int32_t *rowptr, const int64_t *perm, int64_t n, int64_t nnz, int heads,
                         const float *vals, bool edge_major, float *out, const int32_t *long_rows, int64_t n_long,
                         int64_t long_min, hipStream_t s)
{
    GatherArgs ga{rowptr, perm, vals, out, long_rows, n, nnz, long_min, n_long, heads, edge_major ? 1 : 0};
    const dim3 block(256);
    if (heads == 1 || heads == 2 || heads == 4 || heads == 8) {
        const dim3 grid(unsigned(n_long * heads + (n * 16 + 255) / 256));
#define ATTGNN_G_CASE(HV, EMV)                                                                                   \
        if (heads == HV && edge_major == EMV)                                                                    \
            hipLaunchKernelGGL((gather_rowsum_heads_kernel<HV, EMV>), grid, block, 0, s, ga);
        ATTGNN_G_CASE(1, false) ATTGNN_G_CASE(2, false) ATTGNN_G_CASE(4, false) ATTGNN_G_CASE(8, false)
        ATTGNN_G_CASE(1, true) ATTGNN_G_CASE(2, true) ATTGNN_G_CASE(4, true) ATTGNN_G_CASE(8, true)
#undef ATTGNN_G_CASE
        return hipGetLastError();
    }
    if (edge_major) return hipErrorInvalidValue;
    const int64_t threads = n * heads * 16;
    hipLaunchKernelGGL(gather_rowsum_kernel, dim3(unsigned((threads + 255) / 256)), block, 0, s, rowptr, perm, n, nnz,
                       heads, vals, out, long_min);
    if (n_long > 0)
        hipLaunchKernelGGL(gather_rowsum_long_kernel, dim3(unsigned(n_long * heads)), dim3(LONG_THREADS), 0, s, ga);
    return hipGetLastError();
}

// long_min for a row list: rows above ATTGNN_LONG_ROW entries when a list is given.
int64_t long_min_of(const int32_t *long_rows, int64_t n_long)
{
    return (long_rows && n_long > 0) ? int64_t(ATTGNN_LONG_ROW) : INT64_MAX;
}

bool long_args_ok(const int32_t *long_rows, int64_t n_long)
{
    return n_long >= 0 && n_long <= INT32_MAX && (n_long == 0 || long_rows);
}

}  // namespace

extern "C" int attgnn_edge_softmax_fwd(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz,
                                       int64_t heads, const float *s0, const float *s1, int64_t s_head_stride,
                                       int64_t s_row_stride, float slope, float drop_p, uint64_t seed, void *p,
                                       int p_type, float *lse, const int32_t *long_rows, int64_t n_long,
                                       attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0 || heads > 65535 || !float_type(p_type) || drop_p < 0.f || drop_p >= 1.f)
        return ATTGNN_E_ARG;
    if (!long_args_ok(long_rows, n_long)) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !s0 || !s1 || !lse || (nnz > 0 && !col)) return ATTGNN_E_ARG;
    SoftmaxArgs a{};
    a.rowptr = rowptr; a.col = col; a.s0 = s0; a.s1 = s1; a.p = static_cast<uint8_t *>(p); a.lse_out = lse;
    a.n = n; a.nnz = nnz; a.s_hs = s_head_stride; a.s_rs = s_row_stride; a.heads = int(heads); a.p_type = p_type;
    a.slope = slope; a.thresh = drop_threshold(drop_p); a.keep_scale = 1.f / (1.f - drop_p); a.seed = seed;
    a.long_min = long_min_of(long_rows, n_long);
    a.long_rows = long_rows; a.n_long = long_rows ? n_long : 0;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipError_t e;
    if (heads_kernel_ok(heads)) {
        e = launch_fwd_heads(a, node_major(heads, s_head_stride, s_row_stride, s0, s1), s);
    } else {
        const int64_t threads = n * heads * 16;
        hipLaunchKernelGGL(edge_softmax_fwd_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s, a);
        if (a.n_long > 0)
            hipLaunchKernelGGL(softmax_fwd_long_kernel, dim3(unsigned(a.n_long * heads)), dim3(LONG_THREADS), 0, s, a);
        e = hipGetLastError();
    }
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

namespace {

SoftmaxArgs bwd_args(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz, int64_t heads,
                     const float *s0, const float *s1, int64_t s_head_stride, int64_t s_row_stride, const float *lse,
                     const void *dp, int dp_type, float slope, float drop_p, uint64_t seed, float *dz, float *ds0,
                     const int32_t *long_rows, int64_t n_long)
{
    SoftmaxArgs a{};
    a.rowptr = rowptr; a.col = col; a.s0 = s0; a.s1 = s1; a.lse = lse; a.dp = static_cast<const uint8_t *>(dp);
    a.dz = dz; a.ds0 = ds0; a.n = n; a.nnz = nnz; a.s_hs = s_head_stride; a.s_rs = s_row_stride;
    a.heads = int(heads); a.dp_type = dp_type; a.slope = slope; a.thresh = drop_threshold(drop_p);
    a.keep_scale = 1.f / (1.f - drop_p); a.seed = seed;
    a.long_min = long_min_of(long_rows, n_long);
    a.long_rows = long_rows; a.n_long = long_rows ? n_long : 0;
    return a;
}

// Softmax backward over the rows (long rows included).
hipError_t launch_bwd(SoftmaxArgs a, bool edge_major, hipStream_t s)
{
    a.dz_edge_major = edge_major ? 1 : 0;
    const bool nm = node_major(a.heads, a.s_hs, a.s_rs, a.s0, a.s1);
    if (heads_kernel_ok(a.heads))
        return edge_major ? launch_bwd_heads<true>(a, nm, s) : launch_bwd_heads<false>(a, nm, s);
    if (edge_major) return hipErrorInvalidValue;
    const int64_t threads = a.n * a.heads * 16;
    hipLaunchKernelGGL(edge_softmax_bwd_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s, a);
    if (a.n_long > 0)
        hipLaunchKernelGGL(softmax_bwd_long_kernel, dim3(unsigned(a.n_long * a.heads)), dim3(LONG_THREADS), 0, s, a);
    return hipGetLastError();
}

}  // namespace

extern "C" int attgnn_edge_softmax_bwd(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz,
                                       int64_t heads, const float *s0, const float *s1, int64_t s_head_stride,
                                       int64_t s_row_stride, const float *lse, const void *dp, int dp_type,
                                       float slope, float drop_p, uint64_t seed, float *dz, float *ds0,
                                       const int32_t *long_rows, int64_t n_long, attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0 || heads > 65535 || !float_type(dp_type) || drop_p < 0.f || drop_p >= 1.f)
        return ATTGNN_E_ARG;
    if (!long_args_ok(long_rows, n_long)) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !s0 || !s1 || !lse || !ds0 || (nnz > 0 && (!col || !dp || !dz))) return ATTGNN_E_ARG;
    const SoftmaxArgs a = bwd_args(rowptr, col, n, nnz, heads, s0, s1, s_head_stride, s_row_stride, lse, dp, dp_type,
                                   slope, drop_p, seed, dz, ds0, long_rows, n_long);
    const hipError_t e = launch_bwd(a, false, reinterpret_cast<hipStream_t>(stream));
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
                                            const int32_t *t_rowptr, const int64_t *t_perm,
                                            const int32_t *long_rows, int64_t n_long, const int32_t *t_long_rows,
                                            int64_t n_t_long, float *ds0, float *ds1, float *workspace,
                                            int64_t workspace_elems, attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0 || heads > 65535 || !float_type(dp_type) || drop_p < 0.f || drop_p >= 1.f)
        return ATTGNN_E_ARG;
    if (!long_args_ok(long_rows, n_long) || !long_args_ok(t_long_rows, n_t_long)) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !t_rowptr || !s0 || !s1 || !lse || !ds0 || !ds1) return ATTGNN_E_ARG;
    if (nnz > 0 && (!col || !dp || !t_perm || !workspace)) return ATTGNN_E_ARG;
    if (workspace_elems < nnz * heads) return ATTGNN_E_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // dz in the workspace (edge-major for the all-heads kernels), then the column sums
    const bool em = heads_kernel_ok(heads) && (reinterpret_cast<uintptr_t>(workspace) & 15) == 0;
    const SoftmaxArgs a = bwd_args(rowptr, col, n, nnz, heads, s0, s1, s_head_stride, s_row_stride, lse, dp, dp_type,
                                   slope, drop_p, seed, workspace, ds0, long_rows, n_long);
    hipError_t e = launch_bwd(a, em, s);
    if (e == hipSuccess)
        e = launch_gather(t_rowptr, t_perm, n, nnz, int(heads), workspace, em, ds1, t_long_rows, n_t_long,
                          long_min_of(t_long_rows, n_t_long), s);
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_gather_rowsum(const int32_t *rowptr, const int64_t *perm, int64_t n, int64_t nnz,
                                    int64_t heads, const float *vals, float *out, const int32_t *long_rows,
                                    int64_t n_long, attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0 || heads > 65535 || !long_args_ok(long_rows, n_long)) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !out || (nnz > 0 && (!perm || !vals))) return ATTGNN_E_ARG;
    const hipError_t e = launch_gather(rowptr, perm, n, nnz, int(heads), vals, false, out, long_rows, n_long,
                                       long_min_of(long_rows, n_long), reinterpret_cast<hipStream_t>(stream));
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int64_t attgnn_long_rows(const int32_t *rowptr, int64_t n, int32_t *rows)
{
    if (n < 0 || (n > 0 && !rowptr) || n > INT32_MAX) return -1;
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i)
        if (int64_t(rowptr[i + 1]) - rowptr[i] > ATTGNN_LONG_ROW) {
            if (rows) rows[k] = int32_t(i);
            ++k;
        }
    return k;
}

// ---------------------------------------------------------------------------
// Per-head range of the logits z[h, e=(i,j)] = s0[h,i] + s1[h,j] over the CSR's
// edges: the minimum, the maximum and the first edge attaining each.  The
// reference's multi-head GATv2 divides its logits by this range before the
// exponential (end2end_head/gatv2_final/mgat_csr/mgat_conv.py:184-186), a
// per-head temperature; the backward needs the edges that set it.
// Each (value, edge) pair is one 64-bit key -- the value's order-preserving
// bits above the edge index (for the maximum, the index complemented so the
// first edge wins a tie).  A workgroup reduces its keys (lanes, then its waves
// through LDS) into one partial (min, max) per head; a second kernel reduces
// the partials of each head.  Integer min / max, so the result does not depend
// on the order the edges are visited in.  (One 64-bit atomic pair per wave on
// the head's two words instead measured 294 us on 2.48M edges x 4 heads: the
// updates of every XCD serialise on those words.)
namespace {

constexpr int RANGE_HB = 8;   // heads per grid row

__device__ __forceinline__ uint32_t order_bits(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float from_order_bits(uint32_t o)
{
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

struct RangeArgs {
    const int32_t *rowptr, *col;
    const float *s0, *s1;
    int64_t n, s_hs, s_rs;
    unsigned long long *part;   // [heads][ATTGNN_RANGE_BLOCKS][2]: each block's min key, max key
    float *range;               // [heads][3]: min, max, 1/(max - min) (0 when max == min or no edges)
    int32_t *where;             // [heads][4]: row, column of the minimum; row, column of the maximum (-1: none)
    const int32_t *long_rows;   // rows above long_min entries, taken by the first long_blocks workgroups
    int64_t n_long, long_min;
    int heads, blocks, long_blocks;
};

__device__ __forceinline__ unsigned long long umin64(unsigned long long x, unsigned long long y) { return x < y ? x : y; }
__device__ __forceinline__ unsigned long long umax64(unsigned long long x, unsigned long long y) { return x > y ? x : y; }

__device__ __forceinline__ unsigned long long shfl_xor64(unsigned long long v, int o)
{
    const uint32_t lo = __shfl_xor(uint32_t(v), o);
    const uint32_t hi = __shfl_xor(uint32_t(v >> 32), o);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// grid (blocks, ceil(heads / 8)), up to 8 heads per grid row (a column id
// read once for all of them); each block writes its partial keys per head.
__global__ __launch_bounds__(256) void range_kernel(RangeArgs a)
{
    const int h0 = blockIdx.y * RANGE_HB;
    const int nh = min(RANGE_HB, a.heads - h0);
    unsigned long long kmin[RANGE_HB], kmax[RANGE_HB];
#pragma unroll
    for (int k = 0; k < RANGE_HB; ++k) {
        kmin[k] = ~0ull;
        kmax[k] = 0ull;
    }
    // the keys of row i's edges eb, eb + stride, ... below e1, four per lane
    // per iteration with their column ids and column terms loaded before any
    // is used
    auto visit = [&](int64_t i, int32_t eb, int32_t e1, int32_t stride) {
        float r[RANGE_HB];
#pragma unroll
        for (int k = 0; k < RANGE_HB; ++k) r[k] = k < nh ? a.s0[(h0 + k) * a.s_hs + i * a.s_rs] : 0.f;
        for (; eb < e1; eb += 4 * stride) {
            int64_t c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) c[u] = eb + stride * u < e1 ? int64_t(a.col[eb + stride * u]) : -1;
            float t[4][RANGE_HB];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int k = 0; k < RANGE_HB; ++k)
                    t[u][k] = (k < nh && c[u] >= 0) ? a.s1[(h0 + k) * a.s_hs + c[u] * a.s_rs] : 0.f;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (c[u] < 0) continue;
                const uint32_t e = uint32_t(eb + stride * u);
#pragma unroll
                for (int k = 0; k < RANGE_HB; ++k) {
                    if (k < nh) {
                        const float z = r[k] + t[u][k];
                        const unsigned long long hi = static_cast<unsigned long long>(order_bits(z)) << 32;
                        kmin[k] = umin64(kmin[k], hi | e);
                        kmax[k] = umax64(kmax[k], hi | (0xffffffffu - e));
                    }
                }
            }
        }
    };
    if (int(blockIdx.x) < a.long_blocks) {
        // the listed long rows: the whole workgroup walks each row (a hub row
        // of over a thousand entries left to one 16-lane group kept it busy
        // for dozens of dependent load rounds, longer than the rest together)
        for (int64_t li = blockIdx.x; li < a.n_long; li += a.long_blocks) {
            const int64_t i = a.long_rows[li];
            visit(i, a.rowptr[i] + int32_t(threadIdx.x), a.rowptr[i + 1], 256);
        }
    } else {
        // 16-lane groups walk the other rows (grid-stride), a lane taking
        // every 16th edge of its group's row
        const int64_t group = (int64_t(blockIdx.x - a.long_blocks) * blockDim.x + threadIdx.x) >> 4;
        const int64_t groups = (int64_t(gridDim.x - a.long_blocks) * blockDim.x) >> 4;
        for (int64_t i = group; i < a.n; i += groups) {
            const int32_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
            if (e1 - e0 > a.long_min) continue;   // a long-row workgroup takes it
            visit(i, e0 + int32_t(threadIdx.x & 15), e1, 16);
        }
    }
    __shared__ unsigned long long red[4][RANGE_HB][2];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < RANGE_HB; ++k) {
        if (k < nh) {   // wave-uniform: only the launch row's heads are reduced
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                kmin[k] = umin64(kmin[k], shfl_xor64(kmin[k], o));
                kmax[k] = umax64(kmax[k], shfl_xor64(kmax[k], o));
            }
            if ((threadIdx.x & 63) == 0) {
                red[wave][k][0] = kmin[k];
                red[wave][k][1] = kmax[k];
            }
        }
    }
    __syncthreads();
    if (int(threadIdx.x) < nh) {
        const int k = threadIdx.x;
        unsigned long long mn = red[0][k][0], mx = red[0][k][1];
#pragma unroll
        for (int w = 1; w < 4; ++w) {
            mn = umin64(mn, red[w][k][0]);
            mx = umax64(mx, red[w][k][1]);
        }
        unsigned long long *p = a.part + (int64_t(h0 + k) * ATTGNN_RANGE_BLOCKS + blockIdx.x) * 2;
        p[0] = mn;
        p[1] = mx;
    }
}

// Row of edge e: the last i with rowptr[i] <= e (rows are non-decreasing).
__device__ int32_t row_of_edge(const int32_t *rowptr, int64_t n, int32_t e)
{
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (rowptr[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return int32_t(lo);
}

// One workgroup per head: the blocks' partial keys reduced, then the range,
// the temperature and the extreme edges' rows and columns.
__global__ __launch_bounds__(256) void range_finish_kernel(RangeArgs a)
{
    const int h = blockIdx.x;
    unsigned long long mn = ~0ull, mx = 0ull;
    for (int b = threadIdx.x; b < a.blocks; b += blockDim.x) {
        const unsigned long long *p = a.part + (int64_t(h) * ATTGNN_RANGE_BLOCKS + b) * 2;
        mn = umin64(mn, p[0]);
        mx = umax64(mx, p[1]);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        mn = umin64(mn, shfl_xor64(mn, o));
        mx = umax64(mx, shfl_xor64(mx, o));
    }
    __shared__ unsigned long long red[4][2];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = mn;
        red[wave][1] = mx;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    unsigned long long kmin = red[0][0], kmax = red[0][1];
    for (int w = 1; w < int(blockDim.x >> 6); ++w) {
        kmin = umin64(kmin, red[w][0]);
        kmax = umax64(kmax, red[w][1]);
    }
    float *r = a.range + 3 * h;
    int32_t *w = a.where + 4 * h;
    if (kmin == ~0ull) {           // no edges
        r[0] = 0.f; r[1] = 0.f; r[2] = 0.f;
        w[0] = w[1] = w[2] = w[3] = -1;
        return;
    }
    const float zmin = from_order_bits(uint32_t(kmin >> 32)), zmax = from_order_bits(uint32_t(kmax >> 32));
    const int32_t emin = int32_t(uint32_t(kmin)), emax = int32_t(0xffffffffu - uint32_t(kmax));
    r[0] = zmin;
    r[1] = zmax;
    r[2] = zmax > zmin ? 1.f / (zmax - zmin) : 0.f;
    w[0] = row_of_edge(a.rowptr, a.n, emin);
    w[1] = a.col[emin];
    w[2] = row_of_edge(a.rowptr, a.n, emax);
    w[3] = a.col[emax];
}

}  // namespace

extern "C" int64_t attgnn_edge_score_range_workspace_elems(int64_t heads)
{
    return heads > 0 ? heads * ATTGNN_RANGE_BLOCKS * 2 : 0;
}

extern "C" int attgnn_edge_score_range(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz,
                                       int64_t heads, const float *s0, const float *s1, int64_t s_head_stride,
                                       int64_t s_row_stride, unsigned long long *workspace, int64_t workspace_elems,
                                       float *range, int32_t *where, const int32_t *long_rows, int64_t n_long,
                                       attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || nnz > INT32_MAX || heads <= 0 || heads > 256) return ATTGNN_E_ARG;
    if (!long_args_ok(long_rows, n_long)) return ATTGNN_E_ARG;
    if (!range || !where) return ATTGNN_E_ARG;
    if (nnz > 0 && (n == 0 || !rowptr || !col || !s0 || !s1)) return ATTGNN_E_ARG;
    if (nnz > 0 && (!workspace || workspace_elems < attgnn_edge_score_range_workspace_elems(heads)))
        return ATTGNN_E_ARG;
    RangeArgs a{};
    a.rowptr = rowptr; a.col = col; a.s0 = s0; a.s1 = s1; a.n = n; a.s_hs = s_head_stride; a.s_rs = s_row_stride;
    a.part = workspace; a.range = range; a.where = where; a.heads = int(heads);
    a.long_rows = long_rows; a.n_long = n_long; a.long_min = long_min_of(long_rows, n_long);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // up to 1,024 workgroups for the long rows, then a few rows per 16-lane
    // group for the others (enough groups in flight for each row's dependent
    // loads); every block writes its partial slot, so the finish reads only
    // written slots
    a.long_blocks = nnz > 0 ? int(std::min<int64_t>(n_long, 1024)) : 0;
    a.blocks = nnz > 0 ? a.long_blocks + int(std::min<int64_t>(std::max<int64_t>((n + 63) / 64, 1),
                                                                ATTGNN_RANGE_BLOCKS - 1024))
                       : 0;
    if (a.blocks > 0)
        hipLaunchKernelGGL(range_kernel, dim3(unsigned(a.blocks), unsigned((heads + RANGE_HB - 1) / RANGE_HB)),
                           dim3(256), 0, s, a);
    hipLaunchKernelGGL(range_finish_kernel, dim3(unsigned(heads)), dim3(256), 0, s, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

// ---------------------------------------------------------------------------
// Gradient through the min-max logit map (see attgnn_minmax_score_grads in the
// header): U_h partial sums per workgroup, a fixed-order finish per head, then
// one elementwise pass that also applies the four point updates.
namespace {

constexpr int MG_BLOCKS = 256;

struct MinmaxArgs {
    const float *s0, *s1, *range, *ds0s, *ds1s;
    const int32_t *where;
    float *ds0, *ds1, *part;   // part: [heads][MG_BLOCKS] partial U, then [heads] k at the end
    int64_t n, s_hs, s_rs;
    int heads;
};

// block sum in a fixed tree (wave shuffles, then the 4 waves in order)
__device__ __forceinline__ float block_sum_fixed(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(256) void minmax_u_kernel(MinmaxArgs a)
{
    const int h = blockIdx.y;
    float u = 0.f;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < a.n; i += int64_t(gridDim.x) * 256) {
        const int64_t si = h * a.s_hs + i * a.s_rs, di = int64_t(h) * a.n + i;
        u += a.s0[si] * a.ds0s[di] + a.s1[si] * a.ds1s[di];
    }
    u = block_sum_fixed(u);
    if (threadIdx.x == 0) a.part[int64_t(h) * MG_BLOCKS + blockIdx.x] = u;
}

__global__ __launch_bounds__(256) void minmax_k_kernel(MinmaxArgs a)
{
    const int h = blockIdx.x;
    float u = a.part[int64_t(h) * MG_BLOCKS + threadIdx.x];
    u = block_sum_fixed(u);
    if (threadIdx.x == 0) {
        const float c = a.range[3 * h + 2];
        a.part[int64_t(a.heads) * MG_BLOCKS + h] = u * c * c;
    }
}

__global__ __launch_bounds__(256) void minmax_apply_kernel(MinmaxArgs a)
{
    const int h = blockIdx.y;
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= a.n) return;
    const float c = a.range[3 * h + 2], k = a.part[int64_t(a.heads) * MG_BLOCKS + h];
    const int32_t *w = a.where + 4 * h;
    const int64_t di = int64_t(h) * a.n + i;
    float v0 = a.ds0s[di] * c, v1 = a.ds1s[di] * c;
    if (i == w[2]) v0 -= k;   // row of the maximum's edge
    if (i == w[3]) v1 -= k;   // its column
    if (i == w[0]) v0 += k;   // row of the minimum's edge
    if (i == w[1]) v1 += k;   // its column
    a.ds0[di] = v0;
    a.ds1[di] = v1;
}

}  // namespace

extern "C" int64_t attgnn_minmax_grads_workspace_elems(int64_t heads)
{
    return heads > 0 ? heads * (MG_BLOCKS + 1) : 0;
}

extern "C" int attgnn_minmax_score_grads(const float *s0, const float *s1, int64_t n, int64_t heads,
                                         int64_t s_head_stride, int64_t s_row_stride, const float *range,
                                         const int32_t *where, const float *ds0s, const float *ds1s, float *ds0,
                                         float *ds1, float *workspace, int64_t workspace_elems,
                                         attgnn_stream_t stream)
{
    if (n < 0 || heads <= 0 || heads > 65535) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    if (!s0 || !s1 || !range || !where || !ds0s || !ds1s || !ds0 || !ds1 || !workspace) return ATTGNN_E_ARG;
    if (workspace_elems < attgnn_minmax_grads_workspace_elems(heads)) return ATTGNN_E_ARG;
    MinmaxArgs a{};
    a.s0 = s0; a.s1 = s1; a.range = range; a.where = where; a.ds0s = ds0s; a.ds1s = ds1s;
    a.ds0 = ds0; a.ds1 = ds1; a.part = workspace;
    a.n = n; a.s_hs = s_head_stride; a.s_rs = s_row_stride; a.heads = int(heads);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(minmax_u_kernel, dim3(MG_BLOCKS, unsigned(heads)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(minmax_k_kernel, dim3(unsigned(heads)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(minmax_apply_kernel, dim3(unsigned((n + 255) / 256), unsigned(heads)), dim3(256), 0, s, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}
