// GATv2 attention (Brody et al.): logit(i, j) = a . LeakyReLU(x0[i] + x1[j]),
// softmax over the row, h[i] = sum_j p_ij x1[j] -- fused per-edge kernels over
// CSR rows for gfx950.
//
// The reference's "GATv2" (end2end_ori/gatv2_final/mgat_csr/mgat_conv.py:174-196)
// applies the LeakyReLU to the node features before a separable dot product,
// which makes the logit a.X0[i] + a.X1[j] (constant along a softmax row in its
// first term).  This file implements the attention the GATv2 paper defines,
// where the nonlinearity sits between the two endpoints, so the logit is not
// separable and every edge needs a D-wide evaluation:
//   forward  : z[h,e] = sum_d a[h,d] lrelu(x0[i,h,d] + x1[j,h,d]) per edge, the
//              row log-sum-exp, p = exp(z - lse) with dropout;
//   backward : rows pass -- softmax backward dz = p (g - sum p g), dx0[i] =
//              sum_e dz_e a * lrelu'(u_e), and the attention-vector gradient
//              sum_e dz_e lrelu(u_e) reduced in a fixed order;
//              columns pass over the transposed CSR -- dx1[j] = base[j] +
//              sum_{e=(i,j)} dz_e a * lrelu'(u_e).
// Work is cut into segments of at most 32 edges of one row (or of one column,
// for the transposed pass), so a hub row costs many short tasks instead of one
// long chain.  One 16-lane group per (segment, head); lane l owns features
// [l*V, l*V + V) of the head (D = 16 V) or, for other widths, features l,
// l + 16, ...; each lane reads the column id (and dz) of one of 16 edges, and the
// gathers of 8 edges are issued together before their shuffle-reduced dot
// products.  The softmax and its backward run per (row, head) over the stored
// logits (no gathers).  A row with one segment is written directly; rows with
// several sum their segments' partials in order afterwards; the attention-
// vector gradient accumulates per persistent group over strided segments and is
// reduced in a fixed order -- all deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "attgnn.h"
#include "mfma.hpp"

using namespace attgnn;

namespace {

__device__ __forceinline__ float lrelu(float x, float s) { return x > 0.f ? x : x * s; }

__device__ __forceinline__ float group_sum(float v)
{
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) v += __shfl_xor(v, o, 16);
    return v;
}
__device__ __forceinline__ float group_max(float v)
{
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 16));
    return v;
}

__device__ __forceinline__ bool keep_edge(uint64_t seed, int64_t h, int64_t e, uint32_t thresh)
{
    uint64_t x = seed ^ (uint64_t(h) * 0x9E3779B97F4A7C15ull) ^ (uint64_t(e) * 0xC2B2AE3D27D4EB4Full);
    x ^= x >> 31; x *= 0x7FB5D329728EA185ull;
    x ^= x >> 27; x *= 0x81DADEF4BC2DD44Dull;
    x ^= x >> 33;
    return uint32_t(x) >= thresh;
}

__device__ __forceinline__ float ld1(const uint8_t *p, int type, int64_t i)
{
    switch (type) {
    case ATTGNN_F32: return reinterpret_cast<const float *>(p)[i];
    case ATTGNN_F16: return float(reinterpret_cast<const _Float16 *>(p)[i]);
    default: return __uint_as_float(uint32_t(reinterpret_cast<const uint16_t *>(p)[i]) << 16);
    }
}

__device__ __forceinline__ void st1(uint8_t *p, int type, int64_t i, float v)
{
    switch (type) {
    case ATTGNN_F32: reinterpret_cast<float *>(p)[i] = v; break;
    case ATTGNN_F16: reinterpret_cast<_Float16 *>(p)[i] = _Float16(v); break;
    default: reinterpret_cast<uint16_t *>(p)[i] = uint16_t(bf16_bits_rne(v)); break;
    }
}

// Features of one (row, head) slice owned by lane l: V contiguous elements at
// base + l*V (V > 0), or VMAX strided elements l, l + 16, ... below dim (V == 0).
constexpr int VMAX = 8;
constexpr int BATCH = 8;   // edges whose gathers are issued together

template <int V>
struct Span {
    static constexpr int N = V > 0 ? V : VMAX;
    float v[N];

    __device__ __forceinline__ void load(const uint8_t *x, int type, int64_t base, int l, int dim)
    {
        if constexpr (V > 0) {
            const int64_t o = base + int64_t(l) * V;
            if (type == ATTGNN_F32) {
                if constexpr (V % 4 == 0) {
#pragma unroll
                    for (int k = 0; k < V; k += 4) {
                        const f32x4 t = *reinterpret_cast<const f32x4 *>(reinterpret_cast<const float *>(x) + o + k);
                        v[k] = t[0]; v[k + 1] = t[1]; v[k + 2] = t[2]; v[k + 3] = t[3];
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < V; ++k) v[k] = reinterpret_cast<const float *>(x)[o + k];
                }
            } else if constexpr (V % 2 == 0) {
                uint32_t w[V / 2];
                const uint32_t *src = reinterpret_cast<const uint32_t *>(reinterpret_cast<const uint16_t *>(x) + o);
                if constexpr (V == 8) {
                    const u32x4 t = *reinterpret_cast<const u32x4 *>(src);
                    w[0] = t[0]; w[1] = t[1]; w[2] = t[2]; w[3] = t[3];
                } else if constexpr (V == 4) {
                    const u32x2 t = *reinterpret_cast<const u32x2 *>(src);
                    w[0] = t[0]; w[1] = t[1];
                } else {
                    w[0] = src[0];
                }
#pragma unroll
                for (int k = 0; k < V / 2; ++k) {
                    if (type == ATTGNN_BF16) {
                        v[2 * k] = __uint_as_float(w[k] << 16);
                        v[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
                    } else {
                        v[2 * k] = float(__builtin_bit_cast(_Float16, uint16_t(w[k] & 0xffffu)));
                        v[2 * k + 1] = float(__builtin_bit_cast(_Float16, uint16_t(w[k] >> 16)));
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k) v[k] = ld1(x, type, o + k);
            }
        } else {
#pragma unroll
            for (int k = 0; k < VMAX; ++k) {
                const int d = l + 16 * k;
                v[k] = d < dim ? ld1(x, type, base + d) : 0.f;
            }
        }
    }

    __device__ __forceinline__ void load_f32(const float *x, int64_t base, int l, int dim)
    {
        load(reinterpret_cast<const uint8_t *>(x), ATTGNN_F32, base, l, dim);
    }

    __device__ __forceinline__ void store(uint8_t *x, int type, int64_t base, int l, int dim) const
    {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int d = V > 0 ? l * V + k : l + 16 * k;
            if (V > 0 || d < dim) st1(x, type, base + d, v[k]);
        }
    }
};


struct SegTab {
    const int32_t *seg_ptr, *seg_row, *seg_slot, *row_seg;
    int64_t n_seg, n_slots;
};

struct V2Args {
    const int32_t *rowptr, *col;     // CSR of A (rows aggregate over their columns)
    const int32_t *t_col;            // CSR of A^T: the row of A of each entry
    const int64_t *t_perm;           // A edge id of each A^T entry
    SegTab rs, cs;                   // segments of A's rows / A^T's rows
    const uint8_t *x0, *x1;          // [n][heads][dim]
    const float *a;                  // [heads][dim]
    const float *z, *lse, *base;     // logits [heads][nnz], lse [heads][n], base [n][heads][dim] (or null)
    const uint8_t *dp;               // dL/dp after dropout [heads][nnz]
    const float *dzin;
    float *z_out, *lse_out, *dz, *partial_rows, *partial_cols, *partial_da;
    uint8_t *p, *dx0, *dx1;
    int64_t n, nnz, groups;
    int heads, dim, x_type, p_type, dp_type;
    float slope, keep_scale;
    uint32_t thresh;
    uint64_t seed;
};

// ---- forward --------------------------------------------------------------

// z for the edges of one segment of row i.
// 4-head instance: 8 waves per SIMD without spilling when asked (68 VGPRs, 7 waves, otherwise)
template <int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(V == 4 ? 8 : 1))) void gatv2_logits_kernel(V2Args a)
{
    const int64_t gid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (gid >= a.rs.n_seg * a.heads) return;   // whole 16-lane group exits together
    const int64_t sg = gid / a.heads;
    const int h = int(gid - sg * a.heads);
    const int64_t hd = int64_t(a.heads) * a.dim;
    const int64_t i = a.rs.seg_row[sg];
    const int64_t e0 = a.rs.seg_ptr[sg], e1 = a.rs.seg_ptr[sg + 1];
    Span<V> xi, av;
    xi.load(a.x0, a.x_type, i * hd + int64_t(h) * a.dim, l, a.dim);
    av.load_f32(a.a, int64_t(h) * a.dim, l, a.dim);
    for (int64_t b = e0; b < e1; b += 16) {
        const int cnt = e1 - b < 16 ? int(e1 - b) : 16;
        const int cj = l < cnt ? a.col[b + l] : a.col[b];
        float myz = 0.f;
        for (int k0 = 0; k0 < cnt; k0 += BATCH) {
            // the gathers of BATCH edges are in flight together (an edge past
            // the segment's end re-reads the chunk's first row and is not used)
            Span<V> xj[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int64_t j = __shfl(cj, (k0 + u) & 15, 16);
                xj[u].load(a.x1, a.x_type, j * hd + int64_t(h) * a.dim, l, a.dim);
            }
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                float part = 0.f;
#pragma unroll
                for (int q = 0; q < Span<V>::N; ++q) part += av.v[q] * lrelu(xi.v[q] + xj[u].v[q], a.slope);
                const float zk = group_sum(part);
                if (l == k0 + u) myz = zk;
            }
        }
        if (l < cnt) a.z_out[int64_t(h) * a.nnz + b + l] = myz;
    }
}

// Row softmax over stored logits: lse and p (dropout as attgnn_edge_softmax_fwd).
__global__ __launch_bounds__(256) void logit_softmax_fwd_kernel(V2Args a)
{
    const int64_t gid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (gid >= a.n * a.heads) return;
    const int64_t i = gid / a.heads;
    const int h = int(gid - i * a.heads);
    const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    const float *zr = a.z + int64_t(h) * a.nnz;
    float m = -INFINITY, s = 0.f;
    for (int64_t e = e0 + l; e < e1; e += 16) {
        const float v = zr[e];
        const float mn = fmaxf(m, v);
        s = s * __expf(m - mn) + __expf(v - mn);
        m = mn;
    }
    const float M = group_max(m);
    s = (m == -INFINITY) ? 0.f : s * __expf(m - M);
    const float S = group_sum(s);
    const float lse = M + __logf(S);
    if (l == 0) a.lse_out[int64_t(h) * a.n + i] = (e1 > e0) ? lse : 0.f;
    for (int64_t e = e0 + l; e < e1; e += 16) {
        float p = __expf(zr[e] - lse);
        if (a.thresh) p = keep_edge(a.seed, h, e, a.thresh) ? p * a.keep_scale : 0.f;
        st1(a.p, a.p_type, int64_t(h) * a.nnz + e, p);
    }
}

// ---- backward -------------------------------------------------------------

// dz = p (g - sum_k p_k g_k) per row, g = dL/dp before dropout.
__global__ __launch_bounds__(256) void logit_softmax_bwd_kernel(V2Args a)
{
    const int64_t gid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (gid >= a.n * a.heads) return;
    const int64_t i = gid / a.heads;
    const int h = int(gid - i * a.heads);
    const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    const float lse = a.lse[int64_t(h) * a.n + i];
    const int64_t ho = int64_t(h) * a.nnz;
    float acc = 0.f;
    for (int64_t e = e0 + l; e < e1; e += 16) {
        const float p = __expf(a.z[ho + e] - lse);
        float g = ld1(a.dp, a.dp_type, ho + e);
        if (a.thresh) g = keep_edge(a.seed, h, e, a.thresh) ? g * a.keep_scale : 0.f;
        acc += p * g;
    }
    const float S = group_sum(acc);
    for (int64_t e = e0 + l; e < e1; e += 16) {
        const float p = __expf(a.z[ho + e] - lse);
        float g = ld1(a.dp, a.dp_type, ho + e);
        if (a.thresh) g = keep_edge(a.seed, h, e, a.thresh) ? g * a.keep_scale : 0.f;
        a.dz[ho + e] = p * (g - S);
    }
}

// Rows pass.  Persistent group g of head h takes segments g, g + G, ... in
// order: dx0 of the segment's row (direct for a one-segment row, else a
// partial in the row's slot), and its own attention-vector partial.
template <int V>
__global__ __launch_bounds__(256) void gatv2_rows_kernel(V2Args a)
{
    const int64_t gid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (gid >= a.groups * a.heads) return;
    const int64_t g = gid / a.heads;
    const int h = int(gid - g * a.heads);
    const int64_t hd = int64_t(a.heads) * a.dim;
    Span<V> av, da;
    av.load_f32(a.a, int64_t(h) * a.dim, l, a.dim);
#pragma unroll
    for (int q = 0; q < Span<V>::N; ++q) da.v[q] = 0.f;
    for (int64_t sg = g; sg < a.rs.n_seg; sg += a.groups) {
        const int64_t i = a.rs.seg_row[sg];
        const int64_t e0 = a.rs.seg_ptr[sg], e1 = a.rs.seg_ptr[sg + 1];
        Span<V> xi, dx;
        xi.load(a.x0, a.x_type, i * hd + int64_t(h) * a.dim, l, a.dim);
#pragma unroll
        for (int q = 0; q < Span<V>::N; ++q) dx.v[q] = 0.f;
        for (int64_t b = e0; b < e1; b += 16) {
            const int cnt = e1 - b < 16 ? int(e1 - b) : 16;
            const int cj = l < cnt ? a.col[b + l] : a.col[b];
            const float dzl = l < cnt ? a.dzin[int64_t(h) * a.nnz + b + l] : 0.f;
            for (int k0 = 0; k0 < cnt; k0 += BATCH) {
                Span<V> xj[BATCH];
                float dzk[BATCH];
#pragma unroll
                for (int u = 0; u < BATCH; ++u) {
                    const int k = (k0 + u) & 15;
                    const int64_t j = __shfl(cj, k, 16);
                    dzk[u] = k0 + u < cnt ? __shfl(dzl, k, 16) : 0.f;   // dz 0: no contribution
                    xj[u].load(a.x1, a.x_type, j * hd + int64_t(h) * a.dim, l, a.dim);
                }
#pragma unroll
                for (int u = 0; u < BATCH; ++u) {
#pragma unroll
                    for (int q = 0; q < Span<V>::N; ++q) {
                        const float uu = xi.v[q] + xj[u].v[q];
                        dx.v[q] += dzk[u] * av.v[q] * (uu > 0.f ? 1.f : a.slope);
                        da.v[q] += dzk[u] * lrelu(uu, a.slope);
                    }
                }
            }
        }
        const int32_t slot = a.rs.seg_slot[sg];
        if (slot < 0) dx.store(a.dx0, a.x_type, i * hd + int64_t(h) * a.dim, l, a.dim);
        else dx.store(reinterpret_cast<uint8_t *>(a.partial_rows), ATTGNN_F32, (int64_t(slot) * a.heads + h) * a.dim, l,
                      a.dim);
    }
    da.store(reinterpret_cast<uint8_t *>(a.partial_da), ATTGNN_F32, (g * a.heads + h) * int64_t(a.dim), l, a.dim);
}

// Columns pass over segments of A^T's rows: acc = sum dz_e a * lrelu'(u_e);
// a one-segment column writes dx1 = base + acc, else its slot's partial.
template <int V>
__global__ __launch_bounds__(256) void gatv2_cols_kernel(V2Args a)
{
    const int64_t gid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (gid >= a.cs.n_seg * a.heads) return;
    const int64_t sg = gid / a.heads;
    const int h = int(gid - sg * a.heads);
    const int64_t hd = int64_t(a.heads) * a.dim;
    const int64_t j = a.cs.seg_row[sg];
    const int64_t q0 = a.cs.seg_ptr[sg], q1 = a.cs.seg_ptr[sg + 1];
    const int32_t slot = a.cs.seg_slot[sg];
    Span<V> av, xj, acc;
    av.load_f32(a.a, int64_t(h) * a.dim, l, a.dim);
    xj.load(a.x1, a.x_type, j * hd + int64_t(h) * a.dim, l, a.dim);
    if (slot < 0 && a.base) acc.load_f32(a.base, j * hd + int64_t(h) * a.dim, l, a.dim);
    else {
#pragma unroll
        for (int q = 0; q < Span<V>::N; ++q) acc.v[q] = 0.f;
    }
    for (int64_t b = q0; b < q1; b += 16) {
        const int cnt = q1 - b < 16 ? int(q1 - b) : 16;
        int ci = a.t_col[b];
        float dzl = 0.f;
        if (l < cnt) {
            const int64_t e = a.t_perm[b + l];
            ci = a.t_col[b + l];
            dzl = e >= 0 ? a.dzin[int64_t(h) * a.nnz + e] : 0.f;
        }
        for (int k0 = 0; k0 < cnt; k0 += BATCH) {
            Span<V> xi[BATCH];
            float dzk[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int k = (k0 + u) & 15;
                const int64_t i = __shfl(ci, k, 16);
                dzk[u] = k0 + u < cnt ? __shfl(dzl, k, 16) : 0.f;
                xi[u].load(a.x0, a.x_type, i * hd + int64_t(h) * a.dim, l, a.dim);
            }
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
#pragma unroll
                for (int q = 0; q < Span<V>::N; ++q) {
                    const float uu = xi[u].v[q] + xj.v[q];
                    acc.v[q] += dzk[u] * av.v[q] * (uu > 0.f ? 1.f : a.slope);
                }
            }
        }
    }
    if (slot < 0) acc.store(a.dx1, a.x_type, j * hd + int64_t(h) * a.dim, l, a.dim);
    else acc.store(reinterpret_cast<uint8_t *>(a.partial_cols), ATTGNN_F32, (int64_t(slot) * a.heads + h) * a.dim, l,
                   a.dim);
}

// Rows with no segment get base (or zero); rows with several sum their
// partials in segment order (plus base).  One 16-lane group per (row, head).
template <int V>
__global__ __launch_bounds__(256) void seg_combine_kernel(SegTab t, const float *partial, const float *base,
                                                          uint8_t *out, int x_type, int64_t n, int heads, int dim)
{
    const int64_t gid = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (gid >= n * heads) return;
    const int64_t i = gid / heads;
    const int h = int(gid - i * heads);
    const int64_t s0 = t.row_seg[i], s1 = t.row_seg[i + 1];
    if (s1 - s0 == 1) return;   // written by the segment pass
    const int64_t off = (i * heads + h) * int64_t(dim);
    Span<V> acc;
    if (base) acc.load_f32(base, off, l, dim);
    else {
#pragma unroll
        for (int q = 0; q < Span<V>::N; ++q) acc.v[q] = 0.f;
    }
    for (int64_t sg = s0; sg < s1; ++sg) {
        Span<V> part;
        part.load_f32(partial, (int64_t(t.seg_slot[sg]) * heads + h) * dim, l, dim);
#pragma unroll
        for (int q = 0; q < Span<V>::N; ++q) acc.v[q] += part.v[q];
    }
    acc.store(out, x_type, off, l, dim);
}

// Ordered sums of consecutive groups: out[k][e] = sum over y in
// [k*GROUP_SUM, ...) of in[y][e]; with one block left the sums go to da.
constexpr int GROUP_SUM = 128;

__global__ __launch_bounds__(256) void group_sum_kernel(const float *in, int64_t Y, int64_t E, float *out, float *da)
{
    const int64_t blocks = (Y + GROUP_SUM - 1) / GROUP_SUM;
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= blocks * E) return;
    const int64_t k = t / E, e = t - k * E;
    const int64_t y1 = Y < (k + 1) * GROUP_SUM ? Y : (k + 1) * GROUP_SUM;
    float s = 0.f;
    for (int64_t y = k * GROUP_SUM; y < y1; ++y) s += in[y * E + e];
    if (blocks > 1) out[k * E + e] = s;
    else da[e] = s;
}

bool ftype(int t) { return t == ATTGNN_F32 || t == ATTGNN_F16 || t == ATTGNN_BF16; }

uint32_t drop_threshold(float q) { return q <= 0.f ? 0u : uint32_t(std::min(double(q), 0.999999) * 4294967296.0); }

// Lane width: D = 16 V with V in {1, 2, 4, 8}, else the strided path (D <= 128).
int pick_v(int64_t dim)
{
    if (dim == 16 || dim == 32 || dim == 64 || dim == 128) return int(dim / 16);
    return dim <= 16 * VMAX ? 0 : -1;
}

// Persistent groups per head of the rows pass: ~64k (group, head) pairs, at
// most GROUP_SUM^2 per head so the partials reduce in two ordered levels.
int64_t rows_groups(int64_t n_seg, int64_t heads)
{
    const int64_t want = std::max<int64_t>(1, 65536 / std::max<int64_t>(heads, 1));
    return std::max<int64_t>(1, std::min<int64_t>({n_seg, want, int64_t(GROUP_SUM) * GROUP_SUM}));
}

bool aligned(const void *p, int v, int eb) { return p == nullptr || v == 0 || reinterpret_cast<uintptr_t>(p) % (v * eb) == 0; }

SegTab seg_of(const attgnn_segments *s)
{
    SegTab t{};
    t.seg_ptr = s->seg_ptr; t.seg_row = s->seg_row; t.seg_slot = s->seg_slot; t.row_seg = s->row_seg;
    t.n_seg = s->n_seg; t.n_slots = s->n_slots;
    return t;
}

bool seg_ok(const attgnn_segments *s)
{
    return s && s->n_seg >= 0 && s->n_slots >= 0 && s->row_seg &&
           (s->n_seg == 0 || (s->seg_ptr && s->seg_row && s->seg_slot));
}

template <int V> struct Logits { static void run(dim3 g, hipStream_t s, const V2Args &a) { hipLaunchKernelGGL(gatv2_logits_kernel<V>, g, dim3(256), 0, s, a); } };
template <int V> struct Rows { static void run(dim3 g, hipStream_t s, const V2Args &a) { hipLaunchKernelGGL(gatv2_rows_kernel<V>, g, dim3(256), 0, s, a); } };
template <int V> struct Cols { static void run(dim3 g, hipStream_t s, const V2Args &a) { hipLaunchKernelGGL(gatv2_cols_kernel<V>, g, dim3(256), 0, s, a); } };

template <template <int> class K>
void dispatch(int v, dim3 grid, hipStream_t s, const V2Args &a)
{
    switch (v) {
    case 1: K<1>::run(grid, s, a); break;
    case 2: K<2>::run(grid, s, a); break;
    case 4: K<4>::run(grid, s, a); break;
    case 8: K<8>::run(grid, s, a); break;
    default: K<0>::run(grid, s, a); break;
    }
}

void combine(int v, const SegTab &t, const float *partial, const float *base, uint8_t *out, int x_type, int64_t n,
             int heads, int dim, hipStream_t s)
{
    const dim3 grid(unsigned((n * heads * 16 + 255) / 256));
    switch (v) {
    case 1: hipLaunchKernelGGL(seg_combine_kernel<1>, grid, dim3(256), 0, s, t, partial, base, out, x_type, n, heads, dim); break;
    case 2: hipLaunchKernelGGL(seg_combine_kernel<2>, grid, dim3(256), 0, s, t, partial, base, out, x_type, n, heads, dim); break;
    case 4: hipLaunchKernelGGL(seg_combine_kernel<4>, grid, dim3(256), 0, s, t, partial, base, out, x_type, n, heads, dim); break;
    case 8: hipLaunchKernelGGL(seg_combine_kernel<8>, grid, dim3(256), 0, s, t, partial, base, out, x_type, n, heads, dim); break;
    default: hipLaunchKernelGGL(seg_combine_kernel<0>, grid, dim3(256), 0, s, t, partial, base, out, x_type, n, heads, dim); break;
    }
}

}  // namespace

extern "C" int attgnn_gatv2_fwd(const int32_t *rowptr, const int32_t *col, const attgnn_segments *row_segments,
                                int64_t n, int64_t nnz, int64_t heads, int64_t dim, const void *x0, const void *x1,
                                int x_type, const float *att, float slope, float drop_p, uint64_t seed, float *z,
                                float *lse, void *p, int p_type, attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0 || dim <= 0 || !ftype(x_type) || !ftype(p_type) || drop_p < 0.f ||
        drop_p >= 1.f || !seg_ok(row_segments))
        return ATTGNN_E_ARG;
    const int v = pick_v(dim);
    if (v < 0) return ATTGNN_E_UNSUPPORTED;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !x0 || !x1 || !att || !lse || (nnz > 0 && (!col || !z || !p))) return ATTGNN_E_ARG;
    const int eb = x_type == ATTGNN_F32 ? 4 : 2;
    if (!aligned(x0, v, eb) || !aligned(x1, v, eb) || !aligned(att, v, 4)) return ATTGNN_E_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    V2Args a{};
    a.rowptr = rowptr; a.col = col; a.rs = seg_of(row_segments);
    a.x0 = static_cast<const uint8_t *>(x0); a.x1 = static_cast<const uint8_t *>(x1);
    a.a = att; a.z = z; a.z_out = z; a.lse_out = lse; a.p = static_cast<uint8_t *>(p);
    a.n = n; a.nnz = nnz; a.heads = int(heads); a.dim = int(dim); a.x_type = x_type; a.p_type = p_type;
    a.slope = slope; a.thresh = drop_threshold(drop_p); a.keep_scale = 1.f / (1.f - drop_p); a.seed = seed;
    if (a.rs.n_seg > 0)
        dispatch<Logits>(v, dim3(unsigned((a.rs.n_seg * heads * 16 + 255) / 256)), s, a);
    hipLaunchKernelGGL(logit_softmax_fwd_kernel, dim3(unsigned((n * heads * 16 + 255) / 256)), dim3(256), 0, s, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int64_t attgnn_gatv2_bwd_workspace_elems(int64_t heads, int64_t dim, const attgnn_segments *row_segments,
                                                    const attgnn_segments *col_segments)
{
    if (heads <= 0 || dim <= 0 || !row_segments || !col_segments) return -1;
    const int64_t G = rows_groups(row_segments->n_seg, heads);
    const int64_t E = heads * dim;
    return std::max<int64_t>(1, (row_segments->n_slots + col_segments->n_slots + G +
                                 (G + GROUP_SUM - 1) / GROUP_SUM) * E);
}

extern "C" int attgnn_gatv2_bwd(const int32_t *rowptr, const int32_t *col, const attgnn_segments *row_segments,
                                const int32_t *t_col, const int64_t *t_perm, const attgnn_segments *col_segments,
                                int64_t n, int64_t nnz, int64_t heads, int64_t dim, const void *x0, const void *x1,
                                int x_type, const float *att, float slope, float drop_p, uint64_t seed,
                                const float *z, const float *lse, const void *dp, int dp_type, const float *base,
                                float *dz, void *dx0, void *dx1, float *datt, float *workspace,
                                int64_t workspace_elems, attgnn_stream_t stream)
{
    if (n < 0 || nnz < 0 || heads <= 0 || dim <= 0 || !ftype(x_type) || !ftype(dp_type) || drop_p < 0.f ||
        drop_p >= 1.f || !seg_ok(row_segments) || !seg_ok(col_segments))
        return ATTGNN_E_ARG;
    const int v = pick_v(dim);
    if (v < 0) return ATTGNN_E_UNSUPPORTED;
    if (!datt || !workspace ||
        workspace_elems < attgnn_gatv2_bwd_workspace_elems(heads, dim, row_segments, col_segments))
        return ATTGNN_E_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t E = heads * dim;
    if (n == 0) return ATTGNN_OK;
    if (!rowptr || !x0 || !x1 || !att || !lse || !dx0 || !dx1 ||
        (nnz > 0 && (!col || !t_col || !t_perm || !z || !dp || !dz)))
        return ATTGNN_E_ARG;
    const int eb = x_type == ATTGNN_F32 ? 4 : 2;
    if (!aligned(x0, v, eb) || !aligned(x1, v, eb) || !aligned(dx0, v, eb) || !aligned(dx1, v, eb) ||
        !aligned(base, v, 4) || !aligned(att, v, 4))
        return ATTGNN_E_ARG;
    V2Args a{};
    a.rowptr = rowptr; a.col = col; a.t_col = t_col; a.t_perm = t_perm;
    a.rs = seg_of(row_segments); a.cs = seg_of(col_segments);
    a.x0 = static_cast<const uint8_t *>(x0); a.x1 = static_cast<const uint8_t *>(x1); a.a = att;
    a.z = z; a.lse = lse; a.base = base; a.dp = static_cast<const uint8_t *>(dp); a.dz = dz; a.dzin = dz;
    a.partial_rows = workspace;
    a.partial_cols = workspace + a.rs.n_slots * E;
    a.partial_da = a.partial_cols + a.cs.n_slots * E;
    a.dx0 = static_cast<uint8_t *>(dx0); a.dx1 = static_cast<uint8_t *>(dx1);
    a.n = n; a.nnz = nnz; a.heads = int(heads); a.dim = int(dim); a.x_type = x_type; a.dp_type = dp_type;
    a.groups = rows_groups(a.rs.n_seg, heads);
    a.slope = slope; a.thresh = drop_threshold(drop_p); a.keep_scale = 1.f / (1.f - drop_p); a.seed = seed;
    hipLaunchKernelGGL(logit_softmax_bwd_kernel, dim3(unsigned((n * heads * 16 + 255) / 256)), dim3(256), 0, s, a);
    dispatch<Rows>(v, dim3(unsigned((a.groups * heads * 16 + 255) / 256)), s, a);
    if (a.cs.n_seg > 0) dispatch<Cols>(v, dim3(unsigned((a.cs.n_seg * heads * 16 + 255) / 256)), s, a);
    combine(v, a.rs, a.partial_rows, nullptr, a.dx0, x_type, n, int(heads), int(dim), s);
    combine(v, a.cs, a.partial_cols, base, a.dx1, x_type, n, int(heads), int(dim), s);
    // attention-vector gradient: two ordered levels over the groups' partials
    float *level2 = a.partial_da + a.groups * E;
    const float *src = a.partial_da;
    int64_t Y = a.groups;
    while (true) {
        const int64_t blocks = (Y + GROUP_SUM - 1) / GROUP_SUM;
        hipLaunchKernelGGL(group_sum_kernel, dim3(unsigned((blocks * E + 255) / 256)), dim3(256), 0, s, src, Y, E,
                           level2, datt);
        if (blocks == 1) break;
        src = level2;
        Y = blocks;
        level2 = a.partial_da;   // the group partials are consumed; reuse their space
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}
