// Window-block SDDMM and the per-edge attention primitives for gfx950.
//
// Reference counterparts (the reference ships only device headers and Python
// callers for these ops; the kernels themselves are absent):
//   ATT_SDDMM_v{1,3}.sddmm_{fp16,tf32}_h_nnz
//       benchmark-v3/sddmm-benchmark/include/compute_utils.h:142-280,427-530,
//       output_tile.h:5-84 (scatter through the edge-id map)
//   ATT_SDDMM_v{1,3}.att_{fp16,tf32}_h_nnz / attv2_*  (compute_utils.h:1144-1218,
//       output_tile.h:89-202; callers end2end_*/gat*_final/mgat_csr/mgat_conv.py)
//   ATT_SDDMM_v{1,3}.sddmm_{fp16,tf32}_ones_h_nnz     (mgat_conv.py:142-149)
//   ATT_SpMM_v{1,3}.spmm_*_{h_m,m_h}_rowsum           (mgat_conv.py:126-132)
//
// SDDMM: one wave per (part, head).  For each k-block of 32 condensed columns
// the window's rows (A operand) and the 32 gathered column rows (two B tiles)
// are contracted over the feature dimension with MFMA 16x16x32, K = features,
// so every operand is one contiguous 16-byte row load; the 16x32 result is
// scattered to CSR edge order through the edge-id map.
//
// The attention projections a0.X[i] + a1.X[j] are separable, so they are
// computed as two per-node GEMVs (N*H dot products instead of nnz*H) followed
// by a per-edge scatter-add over the window tiles.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "attgnn.h"
#include "mfma.hpp"

using namespace attgnn;

namespace {

__device__ __forceinline__ float load_elt(const uint8_t *p, int type)
{
    switch (type) {
    case ATTGNN_F32: return *reinterpret_cast<const float *>(p);
    case ATTGNN_F16: return float(*reinterpret_cast<const _Float16 *>(p));
    case ATTGNN_BF16: return __uint_as_float(uint32_t(*reinterpret_cast<const uint16_t *>(p)) << 16);
    default: return 0.f;
    }
}

__device__ __forceinline__ void store_elt(uint8_t *p, int type, float v)
{
    switch (type) {
    case ATTGNN_F32: *reinterpret_cast<float *>(p) = v; break;
    case ATTGNN_F16: *reinterpret_cast<_Float16 *>(p) = _Float16(v); break;
    case ATTGNN_BF16: *reinterpret_cast<uint16_t *>(p) = uint16_t(bf16_bits_rne(v)); break;
    default: break;
    }
}

__host__ __device__ inline int type_bytes(int t) { return t == ATTGNN_F32 ? 4 : (t == ATTGNN_F16 || t == ATTGNN_BF16) ? 2 : 0; }

struct SddmmArgs {
    const int32_t *rp, *col, *map, *win;
    const uint8_t *a, *b;     // element (h, i, d) at a[(h*hs + i*rs + d)*eb]
    uint8_t *out;             // out[(h*nnz + e)*ob]
    int64_t parts, nnz, n_rows, dim;
    int64_t a_hs, a_rs, b_hs, b_rs;
    int heads, window, wide_log2, in_type, out_type;
};

// The 16x32 tile C[r][16t+n] = <A[row r], B[col cid[t]]> over all features, for
// one head: lane (g, n) holds C[4g+q][16t+n] in acc[t][q].
template <int KIND, bool VEC>
__device__ __forceinline__ void sddmm_tile(const SddmmArgs &a, int head, const uint8_t *arow, bool row_ok,
                                           const int (&cid)[2], int g, f32x4 (&acc)[2])
{
    constexpr int EB = KIND >= 2 ? 4 : 2;
    acc[0] = f32x4{0.f, 0.f, 0.f, 0.f};
    acc[1] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int64_t d0 = 0; d0 < a.dim; d0 += 32) {
        const int64_t d = d0 + 8 * g;
        // 8 features d..d+7 of the A row and of the two B rows
        float av[8], bv[2][8];
#pragma unroll
        for (int i = 0; i < 8; ++i) { av[i] = 0.f; bv[0][i] = 0.f; bv[1][i] = 0.f; }
        uint32_t araw[4] = {0, 0, 0, 0}, braw[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
        if constexpr (VEC && EB == 2) {
            if (d < a.dim) {
                if (row_ok) {
                    const u32x4 v = *reinterpret_cast<const u32x4 *>(arow + d * 2);
                    araw[0] = v[0]; araw[1] = v[1]; araw[2] = v[2]; araw[3] = v[3];
                }
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    if (cid[t] >= 0) {
                        const u32x4 v = *reinterpret_cast<const u32x4 *>(
                            a.b + (int64_t(head) * a.b_hs + int64_t(cid[t]) * a.b_rs + d) * 2);
                        braw[t][0] = v[0]; braw[t][1] = v[1]; braw[t][2] = v[2]; braw[t][3] = v[3];
                    }
            }
        } else if constexpr (VEC && EB == 4) {
            if (d < a.dim) {
                if (row_ok)
#pragma unroll
                    for (int i = 0; i < 8; ++i) av[i] = reinterpret_cast<const float *>(arow)[d + i];
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    if (cid[t] >= 0) {
                        const float *br = reinterpret_cast<const float *>(
                            a.b + (int64_t(head) * a.b_hs + int64_t(cid[t]) * a.b_rs) * 4);
#pragma unroll
                        for (int i = 0; i < 8; ++i) bv[t][i] = br[d + i];
                    }
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (d + i < a.dim) {
                    if (row_ok) av[i] = load_elt(arow + (d + i) * EB, a.in_type);
#pragma unroll
                    for (int t = 0; t < 2; ++t)
                        if (cid[t] >= 0)
                            bv[t][i] = load_elt(a.b + (int64_t(head) * a.b_hs + int64_t(cid[t]) * a.b_rs + d + i) * EB,
                                                a.in_type);
                }
            }
            if constexpr (EB == 2) {  // repack to 16-bit operands
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    auto enc = [](float x) -> uint32_t {
                        if constexpr (KIND == 0) return __builtin_bit_cast(uint16_t, _Float16(x));
                        else return bf16_bits_rne(x);
                    };
                    araw[q] = enc(av[2 * q]) | enc(av[2 * q + 1]) << 16;
#pragma unroll
                    for (int t = 0; t < 2; ++t) braw[t][q] = enc(bv[t][2 * q]) | enc(bv[t][2 * q + 1]) << 16;
                }
            }
        }
        if constexpr (KIND == 0 || KIND == 1) {
            const u32x4 A = {araw[0], araw[1], araw[2], araw[3]};
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const u32x4 B = {braw[t][0], braw[t][1], braw[t][2], braw[t][3]};
                acc[t] = KIND == 0 ? mfma_f16(A, B, acc[t]) : mfma_bf16(A, B, acc[t]);
            }
        } else if constexpr (KIND == 3) {
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int i = 0; i < 8; ++i) acc[t] = mfma_f32(av[i], bv[t][i], acc[t]);
        } else {
            uint32_t ah[8], al[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) bf16_split(av[i], ah[i], al[i]);
            const u32x4 AH = {ah[0] | ah[1] << 16, ah[2] | ah[3] << 16, ah[4] | ah[5] << 16, ah[6] | ah[7] << 16};
            const u32x4 AL = {al[0] | al[1] << 16, al[2] | al[3] << 16, al[4] | al[5] << 16, al[6] | al[7] << 16};
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                uint32_t bh[8], bl[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) bf16_split(bv[t][i], bh[i], bl[i]);
                const u32x4 BH = {bh[0] | bh[1] << 16, bh[2] | bh[3] << 16, bh[4] | bh[5] << 16, bh[6] | bh[7] << 16};
                const u32x4 BL = {bl[0] | bl[1] << 16, bl[2] | bl[3] << 16, bl[4] | bl[5] << 16, bl[6] | bl[7] << 16};
                acc[t] = mfma_bf16(AL, BH, acc[t]);
                acc[t] = mfma_bf16(AH, BL, acc[t]);
                acc[t] = mfma_bf16(AH, BH, acc[t]);
            }
        }
    }
}

// VEC: rows are 16-byte aligned runs of 8 (16-bit) or 4 (fp32) elements and
// dim is a multiple of 8, so each lane loads its 8 features with vector loads.
// KIND: 0 = f16, 1 = bf16, 2 = fp32 via bf16x3, 3 = fp32 MFMA.
template <int KIND, bool VEC>
__global__ __launch_bounds__(256) void sddmm_block_kernel(SddmmArgs a)
{
    constexpr int EB = KIND >= 2 ? 4 : 2;
    const int lane = threadIdx.x & 63;
    // wave index in a scalar register: the task stays wave-uniform
    const int64_t task = int64_t(blockIdx.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (task >= a.parts * a.heads) return;
    const int64_t part = task / a.heads;
    const int head = int(task - part * a.heads);
    const int c_begin = a.rp[2 * part];
    const int plen = a.rp[2 * part + 2] - c_begin;
    const int nreal = a.rp[2 * part + 1] - c_begin;
    const int g = lane >> 4, n = lane & 15;
    const int W = a.window, wide = 1 << a.wide_log2;
    const int64_t row = int64_t(a.win[part]) * W + n;  // A-operand row of this lane
    const bool row_ok = n < W && row < a.n_rows;
    const uint8_t *arow = a.a + (int64_t(head) * a.a_hs + row * a.a_rs) * EB;

    for (int kb = 0; kb < nreal; kb += 32) {
        f32x4 acc[2];
        int cid[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int k = kb + 16 * t + n;
            cid[t] = k < nreal ? a.col[c_begin + k] : -1;
        }
        sddmm_tile<KIND, VEC>(a, head, arow, row_ok, cid, g, acc);
        // scatter C[4g+q][16t+n] to its edge
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int k = kb + 16 * t + n;
            if (k >= nreal) continue;
            const int64_t cc = int64_t(c_begin) + k;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = 4 * g + q;
                if (r >= W) break;
                const int e = a.map[(cc >> a.wide_log2) * W * wide + int64_t(r) * wide + (cc & (wide - 1))];
                if (e >= 0) store_elt(a.out + (int64_t(head) * a.nnz + e) * type_bytes(a.out_type), a.out_type, acc[t][q]);
            }
        }
    }
}

// The staged 16 x 32 product tile has rows 36 floats apart: the MFMA output
// writes of the four 16-lane groups (rows 4g+q, one column per lane) then fall
// in distinct banks (a 32-float pitch put every group on the same 16 banks).
// Tile position pos = row * 32 + column (the edge list's) sits at s_at(pos).
constexpr int S_PITCH = 36;
__device__ __forceinline__ int s_at(int pos) { return (pos >> 5) * S_PITCH + (pos & 31); }

// SDDMM over compact edge lists: one wave per part, every head in the same
// wave.  Per 32-column block the wave computes the 16x32 tile with MFMA (all
// rows of the window against the block's 32 gathered columns), stages it in a
// wave-private 2.25 KB LDS tile and writes only the block's real edges, listed as
// (packed position r*32+c, edge id) -- the dense edge-id map is never read.
// Column ids and the edge list are loaded once per block for all heads.
template <int KIND, bool VEC>
__global__ __launch_bounds__(256) void sddmm_edges_kernel(SddmmArgs a, const int64_t *bbase, const int32_t *eids,
                                                          const uint16_t *epos)
{
    constexpr int EB = KIND >= 2 ? 4 : 2;
    __shared__ float tile[4][16 * S_PITCH];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform part
    const int64_t part = int64_t(blockIdx.x) * 4 + wave;
    if (part >= a.parts) return;
    const int c_begin = a.rp[2 * part];
    const int nreal = a.rp[2 * part + 1] - c_begin;
    const int g = lane >> 4, n = lane & 15;
    const int W = a.window;
    const int64_t row = int64_t(a.win[part]) * W + n;
    const bool row_ok = n < W && row < a.n_rows;
    float *S = tile[wave];
    const int ob = type_bytes(a.out_type);

    for (int kb = 0; kb < nreal; kb += 32) {
        const int64_t blk = (int64_t(c_begin) + kb) >> 5;
        const int64_t e0 = bbase[blk];
        const int cnt = int(bbase[blk + 1] - e0);
        if (cnt == 0) continue;
        int cid[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int k = kb + 16 * t + n;
            cid[t] = k < nreal ? a.col[c_begin + k] : -1;
        }
        // this lane's first edge of the block (most blocks have < 64 edges)
        int my_pos = -1, my_e = -1;
        if (lane < cnt) {
            my_pos = epos[e0 + lane];
            my_e = eids[e0 + lane];
        }
        for (int h = 0; h < a.heads; ++h) {
            const uint8_t *arow = a.a + (int64_t(h) * a.a_hs + row * a.a_rs) * EB;
            f32x4 acc[2];
            sddmm_tile<KIND, VEC>(a, h, arow, row_ok, cid, g, acc);
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) S[(4 * g + q) * S_PITCH + 16 * t + n] = acc[t][q];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (my_e >= 0) store_elt(a.out + (int64_t(h) * a.nnz + my_e) * ob, a.out_type, S[s_at(my_pos)]);
            for (int j = 64 + lane; j < cnt; j += 64) {
                const int e = eids[e0 + j];
                store_elt(a.out + (int64_t(h) * a.nnz + e) * ob, a.out_type, S[s_at(epos[e0 + j])]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

// 16-bit inputs with 16-byte aligned rows and dim <= 32*NDC: the gathers of a
// block's (head, 32-feature chunk) pairs are issued PAIRS at a time before
// their MFMAs (whole heads per batch), and the next block's column ids and
// edge list load while this block computes.  Same tiles, same results as
// sddmm_edges_kernel.
constexpr int PAIRS = 8;

template <int KIND, int NDC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void sddmm_edges_batched_kernel(SddmmArgs a, const int64_t *bbase,
                                                                  const int32_t *eids, const uint16_t *epos)
{
    constexpr int HPB = PAIRS / NDC;   // heads per batch
    __shared__ float tile[4][16 * S_PITCH];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform part
    const int64_t part = int64_t(blockIdx.x) * 4 + wave;
    if (part >= a.parts) return;
    const int c_begin = a.rp[2 * part];
    const int nreal = a.rp[2 * part + 1] - c_begin;
    const int g = lane >> 4, n = lane & 15;
    const int W = a.window;
    const int64_t row = int64_t(a.win[part]) * W + n;
    const bool row_ok = n < W && row < a.n_rows;
    float *S = tile[wave];
    const int ob = type_bytes(a.out_type);

    // block metadata: column ids of the lane's two columns, the block's edge
    // range and this lane's first edge
    int cid[2] = {-1, -1}, cnt = 0, pos = -1, eid = -1;
    int64_t e0 = 0;
    auto meta = [&](int kb, int (&c)[2], int64_t &b0, int &bc, int &bp, int &be) {
        const int64_t blk = (int64_t(c_begin) + kb) >> 5;
        b0 = bbase[blk];
        bc = int(bbase[blk + 1] - b0);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int k = kb + 16 * t + n;
            c[t] = k < nreal ? a.col[c_begin + k] : -1;
        }
        bp = -1;
        be = -1;
        if (lane < bc) {
            bp = epos[b0 + lane];
            be = eids[b0 + lane];
        }
    };
    // The window rows (A operand) are the same for every k-block: when all
    // heads fit in one batch they are loaded once for the part.
    const bool one_batch = a.heads <= HPB;
    u32x4 ar[PAIRS];
    auto load_a = [&](int h0, int hn) {
#pragma unroll
        for (int p = 0; p < PAIRS; ++p) {
            const int hl = p / NDC;
            const int64_t d = (p % NDC) * 32 + 8 * g;
            ar[p] = u32x4{0, 0, 0, 0};
            if (hl < hn && d < a.dim && row_ok)
                ar[p] = *reinterpret_cast<const u32x4 *>(a.a + ((h0 + hl) * a.a_hs + row * a.a_rs + d) * 2);
        }
    };
    if (one_batch && nreal > 0) load_a(0, a.heads);
    if (nreal > 0) meta(0, cid, e0, cnt, pos, eid);
    for (int kb = 0; kb < nreal; kb += 32) {
        int ncid[2] = {-1, -1}, ncnt = 0, npos = -1, neid = -1;
        int64_t ne0 = 0;
        if (kb + 32 < nreal) meta(kb + 32, ncid, ne0, ncnt, npos, neid);
        if (cnt > 0) {
            for (int h0 = 0; h0 < a.heads; h0 += HPB) {
                const int hn = a.heads - h0 < HPB ? a.heads - h0 : HPB;
                if (!one_batch) load_a(h0, hn);
                u32x4 br[PAIRS][2];
#pragma unroll
                for (int p = 0; p < PAIRS; ++p) {
                    const int hl = p / NDC;
                    const int64_t d = (p % NDC) * 32 + 8 * g;
                    const int64_t h = h0 + hl;
                    br[p][0] = br[p][1] = u32x4{0, 0, 0, 0};
                    if (hl < hn && d < a.dim) {
#pragma unroll
                        for (int t = 0; t < 2; ++t)
                            if (cid[t] >= 0)
                                br[p][t] = *reinterpret_cast<const u32x4 *>(
                                    a.b + (h * a.b_hs + int64_t(cid[t]) * a.b_rs + d) * 2);
                    }
                }
                f32x4 acc[HPB][2];
#pragma unroll
                for (int hl = 0; hl < HPB; ++hl) acc[hl][0] = acc[hl][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int p = 0; p < PAIRS; ++p)
#pragma unroll
                    for (int t = 0; t < 2; ++t)
                        acc[p / NDC][t] = KIND == 0 ? mfma_f16(ar[p], br[p][t], acc[p / NDC][t])
                                                    : mfma_bf16(ar[p], br[p][t], acc[p / NDC][t]);
#pragma unroll
                for (int hl = 0; hl < HPB; ++hl) {
                    if (hl >= hn) break;
                    const int64_t h = h0 + hl;
#pragma unroll
                    for (int t = 0; t < 2; ++t)
#pragma unroll
                        for (int q = 0; q < 4; ++q) S[(4 * g + q) * S_PITCH + 16 * t + n] = acc[hl][t][q];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (eid >= 0) store_elt(a.out + (h * a.nnz + eid) * ob, a.out_type, S[s_at(pos)]);
                    for (int j = 64 + lane; j < cnt; j += 64) {
                        const int e = eids[e0 + j];
                        store_elt(a.out + (h * a.nnz + e) * ob, a.out_type, S[s_at(epos[e0 + j])]);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
        }
        cid[0] = ncid[0]; cid[1] = ncid[1];
        e0 = ne0; cnt = ncnt; pos = npos; eid = neid;
    }
}

// s[h*s_hs + i*s_rs] = sum_d X[h, i, d] * vec[h*dim + d]   (fp32 accumulate)
// Generic form: 16 lanes per (h, i), one element per lane per step.
__global__ __launch_bounds__(256) void node_score_kernel(const uint8_t *x, int x_type, int64_t x_hs,
                                                         int64_t x_rs, const uint8_t *vec, int v_type,
                                                         int64_t heads, int64_t n_rows, int64_t dim,
                                                         float *s, int64_t s_hs, int64_t s_rs)
{
    const int64_t idx = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    const bool ok = idx < heads * n_rows;
    float acc = 0.f;
    int64_t h = 0, i = 0;
    if (ok) {
        h = idx / n_rows;
        i = idx - h * n_rows;
        const int xe = type_bytes(x_type), ve = type_bytes(v_type);
        const uint8_t *xr = x + (h * x_hs + i * x_rs) * xe;
        const uint8_t *vr = vec + h * dim * ve;
        for (int64_t d = l; d < dim; d += 16) acc += load_elt(xr + d * xe, x_type) * load_elt(vr + d * ve, v_type);
    }
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 16);
    if (ok && l == 0) s[h * s_hs + i * s_rs] = acc;
}

// Vector form: 16 lanes per node i with all H heads.  The node's H x dim
// values are cut into 16-byte chunks (CE elements); lane l takes chunks
// l, l+16, ...  Each chunk's dot product is summed in element order and added
// to its head's partial, then the 16 partials of each head are reduced.
// NV = 2: a second vector (vec2 -> s2) from the same loads of X, each score
// summed in the same order as alone.
template <int EB, int H, int NV>
__global__ __launch_bounds__(256) void node_score_vec_kernel(const uint8_t *x, int x_type, int64_t x_hs,
                                                             int64_t x_rs, const uint8_t *vec, int v_type,
                                                             int64_t n_rows, int64_t dim, float *s,
                                                             int64_t s_hs, int64_t s_rs, const uint8_t *vec2,
                                                             float *s2)
{
    constexpr int CE = 16 / EB;
    const int64_t i = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    if (i >= n_rows) return;          // whole 16-lane group exits together
    const int64_t cph = dim / CE, chunks = cph * H;
    const bool same = v_type == x_type;
    float acc[H], acc2[H];
#pragma unroll
    for (int h = 0; h < H; ++h) acc[h] = acc2[h] = 0.f;
    for (int64_t c = l; c < chunks; c += 16) {
        const int h = int(c / cph);
        const int64_t d0 = (c - int64_t(h) * cph) * CE;
        const u32x4 xv = *reinterpret_cast<const u32x4 *>(x + (int64_t(h) * x_hs + i * x_rs + d0) * EB);
        float xf[CE];
        if constexpr (EB == 4) {
#pragma unroll
            for (int k = 0; k < 4; ++k) xf[k] = __uint_as_float(xv[k]);
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t b = (xv[k >> 1] >> (16 * (k & 1))) & 0xffffu;
                xf[k] = x_type == ATTGNN_F16 ? float(__builtin_bit_cast(_Float16, uint16_t(b))) : __uint_as_float(b << 16);
            }
        }
        auto chunk_dot = [&](const uint8_t *vbase) {
            float vf[CE];
            const uint8_t *vp = vbase + (int64_t(h) * dim + d0) * type_bytes(v_type);
            if (same) {
                const u32x4 vv = *reinterpret_cast<const u32x4 *>(vp);
                if constexpr (EB == 4) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) vf[k] = __uint_as_float(vv[k]);
                } else {
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const uint32_t b = (vv[k >> 1] >> (16 * (k & 1))) & 0xffffu;
                        vf[k] = v_type == ATTGNN_F16 ? float(__builtin_bit_cast(_Float16, uint16_t(b))) : __uint_as_float(b << 16);
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < CE; ++k) vf[k] = load_elt(vp + k * type_bytes(v_type), v_type);
            }
            float p = 0.f;
#pragma unroll
            for (int k = 0; k < CE; ++k) p += xf[k] * vf[k];
            return p;
        };
        const float p = chunk_dot(vec);
#pragma unroll
        for (int hh = 0; hh < H; ++hh)
            if (hh == h) acc[hh] += p;
        if constexpr (NV == 2) {
            const float p2 = chunk_dot(vec2);
#pragma unroll
            for (int hh = 0; hh < H; ++hh)
                if (hh == h) acc2[hh] += p2;
        }
    }
#pragma unroll
    for (int h = 0; h < H; ++h) {
        float v = acc[h];
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 16);
        if (l == 0) s[h * s_hs + i * s_rs] = v;
        if constexpr (NV == 2) {
            float v2 = acc2[h];
#pragma unroll
            for (int off = 8; off >= 1; off >>= 1) v2 += __shfl_xor(v2, off, 16);
            if (l == 0) s2[h * s_hs + i * s_rs] = v2;
        }
    }
}

struct EdgeArgs {
    const int32_t *rp, *col, *map, *win;
    const float *s0, *s1;       // node terms, element (h, i) at s[h*hs + i*rs]; s1 may be null
    const uint8_t *g0;          // alternative typed source for s0 (ones op), or null
    int g0_type;
    int64_t s0_hs, s0_rs, s1_hs, s1_rs;
    uint8_t *out;
    int64_t parts, nnz, n_rows;
    int heads, window, wide_log2, out_type;
};

// out[h*nnz + e] = s0[h, row(e)] + s1[h, col(e)] over every tile slot of a part.
__global__ __launch_bounds__(256) void edge_from_nodes_kernel(EdgeArgs a)
{
    const int64_t part = blockIdx.x;
    const int W = a.window, wide = 1 << a.wide_log2;
    const int64_t c_begin = a.rp[2 * part];
    const int64_t plen = a.rp[2 * part + 2] - c_begin;
    const int64_t w = a.win[part];
    const int ob = type_bytes(a.out_type);
    for (int64_t s = threadIdx.x; s < plen * W; s += blockDim.x) {
        const int64_t idx = c_begin * W + s;
        const int e = a.map[idx];
        if (e < 0) continue;
        const int64_t within = idx & (int64_t(W) * wide - 1);   // W*wide is a power of two
        const int r = int(within >> a.wide_log2);
        const int64_t cc = (idx / (int64_t(W) * wide)) * wide + (within & (wide - 1));
        const int64_t row = w * W + r;
        const int64_t c = a.col[cc];
        for (int h = 0; h < a.heads; ++h) {
            float v = 0.f;
            if (a.g0) v = row < a.n_rows ? load_elt(a.g0 + (h * a.s0_hs + row * a.s0_rs) * type_bytes(a.g0_type), a.g0_type) : 0.f;
            else if (a.s0) v = row < a.n_rows ? a.s0[h * a.s0_hs + row * a.s0_rs] : 0.f;
            if (a.s1) v += a.s1[h * a.s1_hs + c * a.s1_rs];
            store_elt(a.out + (int64_t(h) * a.nnz + e) * ob, a.out_type, v);
        }
    }
}

// rowsum[h, r] = sum over the row's tile slots of val[h*nnz + map[slot]].
// One workgroup per part; the first part of each window sums all of the
// window's parts in a fixed order (deterministic).
__global__ __launch_bounds__(256) void rowsum_map_kernel(const int32_t *rp, const int32_t *map,
                                                         const int32_t *win, int64_t parts,
                                                         const uint8_t *val, int val_type, int64_t nnz,
                                                         int heads, int window, int wide_log2,
                                                         float *out, int64_t out_hs, int64_t out_rs,
                                                         int64_t m_rows)
{
    __shared__ float red[256];
    const int64_t p = blockIdx.x;
    const int w = win[p];
    if (p > 0 && win[p - 1] == w) return;
    int64_t last = p;
    while (last + 1 < parts && win[last + 1] == w) ++last;
    const int W = window, wide = 1 << wide_log2;
    const int64_t c_lo = rp[2 * p], c_hi = rp[2 * last + 2];
    const int lanes = 256 / W;                 // threads per row
    const int r = threadIdx.x / lanes, l = threadIdx.x - r * lanes;
    const int vb = type_bytes(val_type);
    for (int h = 0; h < heads; ++h) {
        float acc = 0.f;
        for (int64_t cc = c_lo + l; cc < c_hi; cc += lanes) {
            const int e = map[(cc >> wide_log2) * W * wide + int64_t(r) * wide + (cc & (wide - 1))];
            if (e >= 0) acc += load_elt(val + (int64_t(h) * nnz + e) * vb, val_type);
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int off = lanes / 2; off >= 1; off >>= 1) {
            if (l < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        const int64_t row = int64_t(w) * W + r;
        if (l == 0 && row < m_rows) out[h * out_hs + row * out_rs] = red[threadIdx.x];
        __syncthreads();
    }
}


// Backward of the per-node attention terms s0 = Xs.a0, s1 = Xt.a1 over
// [n][heads][dim] features (C = heads*dim columns per row), fused with the
// aggregation gradient:
//   dxs[i,c] = ds0[h,i] a0[c]                         (dxs != null)
//   dxt[i,c] = base[i,c] + ds1[h,i] a1[c] (+ ds0[h,i] a0[c] when dxs == null)
//   partial[y][0][c] = sum_{i in chunk y} ds0[h,i] xs[i,c],  [y][1][c] likewise
// with h = c / dim.  A 256-thread workgroup is 64 column lanes (V columns
// each, one head) x 4 row lanes; rows are walked in order per lane and the 4
// row lanes are combined in order, so the partials are deterministic.
struct NodeBwdArgs {
    const uint8_t *xs, *xt;
    const float *a0, *a1, *ds0, *ds1, *base;
    uint8_t *dxs, *dxt;
    float *partial;
    int64_t n, heads, dim, C;
    int x_type;
};

template <int V>
__device__ __forceinline__ void load_vec(const uint8_t *p, int type, float (&v)[V])
{
    if constexpr (V == 4) {
        if (type == ATTGNN_F32) {
            const f32x4 x = *reinterpret_cast<const f32x4 *>(p);
            v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; v[3] = x[3];
            return;
        }
        const u32x2 x = *reinterpret_cast<const u32x2 *>(p);
        if (type == ATTGNN_BF16) {
            v[0] = __uint_as_float(x[0] << 16); v[1] = __uint_as_float(x[0] & 0xffff0000u);
            v[2] = __uint_as_float(x[1] << 16); v[3] = __uint_as_float(x[1] & 0xffff0000u);
        } else {
            const _Float16 *h = reinterpret_cast<const _Float16 *>(&x);
            v[0] = float(h[0]); v[1] = float(h[1]); v[2] = float(h[2]); v[3] = float(h[3]);
        }
    } else {
        const int eb = type_bytes(type);
#pragma unroll
        for (int k = 0; k < V; ++k) v[k] = load_elt(p + k * eb, type);
    }
}

template <int V>
__device__ __forceinline__ void store_vec(uint8_t *p, int type, const float (&v)[V])
{
    if constexpr (V == 4) {
        if (type == ATTGNN_F32) {
            *reinterpret_cast<f32x4 *>(p) = f32x4{v[0], v[1], v[2], v[3]};
            return;
        }
        uint32_t h[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            h[k] = type == ATTGNN_BF16 ? bf16_bits_rne(v[k]) : uint32_t(__builtin_bit_cast(uint16_t, _Float16(v[k])));
        *reinterpret_cast<u32x2 *>(p) = u32x2{h[0] | h[1] << 16, h[2] | h[3] << 16};
    } else {
        const int eb = type_bytes(type);
#pragma unroll
        for (int k = 0; k < V; ++k) store_elt(p + k * eb, type, v[k]);
    }
}

template <int V>
__global__ __launch_bounds__(256) void attn_node_bwd_kernel(NodeBwdArgs a)
{
    __shared__ float red[2][4][64 * V];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t c0 = (int64_t(blockIdx.x) * 64 + tx) * V;
    const bool active = c0 < a.C;
    const int64_t h = active ? c0 / a.dim : 0;
    const int eb = type_bytes(a.x_type);
    float av0[V], av1[V], acc0[V], acc1[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        av0[k] = active ? a.a0[c0 + k] : 0.f;
        av1[k] = active ? a.a1[c0 + k] : 0.f;
        acc0[k] = acc1[k] = 0.f;
    }
    const int64_t per = (a.n + gridDim.y - 1) / gridDim.y;
    const int64_t r0 = int64_t(blockIdx.y) * per, r1 = min(a.n, r0 + per);
    if (active) {
        const bool same = a.xs == a.xt;
        for (int64_t i = r0 + ty; i < r1; i += 4) {
            const float g0 = a.ds0[h * a.n + i], g1 = a.ds1[h * a.n + i];
            const int64_t off = i * a.C + c0;
            float xs[V], xt[V], b[V], o[V];
            load_vec<V>(a.xs + off * eb, a.x_type, xs);
            if (same) {
#pragma unroll
                for (int k = 0; k < V; ++k) xt[k] = xs[k];
            } else {
                load_vec<V>(a.xt + off * eb, a.x_type, xt);
            }
            if (a.base) load_vec<V>(reinterpret_cast<const uint8_t *>(a.base + off), ATTGNN_F32, b);
            else {
#pragma unroll
                for (int k = 0; k < V; ++k) b[k] = 0.f;
            }
#pragma unroll
            for (int k = 0; k < V; ++k) {
                acc0[k] += g0 * xs[k];
                acc1[k] += g1 * xt[k];
            }
            if (a.dxs) {
#pragma unroll
                for (int k = 0; k < V; ++k) o[k] = g0 * av0[k];
                store_vec<V>(a.dxs + off * eb, a.x_type, o);
#pragma unroll
                for (int k = 0; k < V; ++k) o[k] = b[k] + g1 * av1[k];
            } else {
#pragma unroll
                for (int k = 0; k < V; ++k) o[k] = b[k] + g1 * av1[k] + g0 * av0[k];
            }
            store_vec<V>(a.dxt + off * eb, a.x_type, o);
        }
    }
#pragma unroll
    for (int k = 0; k < V; ++k) {
        red[0][ty][tx * V + k] = acc0[k];
        red[1][ty][tx * V + k] = acc1[k];
    }
    __syncthreads();
    // 256 threads combine 2 x 64V values over the 4 row lanes, in order
    for (int e = threadIdx.x; e < 2 * 64 * V; e += 256) {
        const int which = e / (64 * V), c = e - which * 64 * V;
        const int64_t col = int64_t(blockIdx.x) * 64 * V + c;
        if (col >= a.C) continue;
        const float sum = ((red[which][0][c] + red[which][1][c]) + red[which][2][c]) + red[which][3][c];
        a.partial[(int64_t(blockIdx.y) * 2 + which) * a.C + col] = sum;
    }
}

// Sum of consecutive chunk partials, in order: out[g][e] = sum over
// y in [g*G, min(g*G + G, Y)) of in[y][e] (E = 2C values per chunk).  When one
// group remains the sums go straight to da0 (e < C) / da1 (e >= C).
constexpr int CHUNK_GROUP = 64;

__global__ __launch_bounds__(256) void chunk_sum_kernel(const float *in, int64_t Y, int64_t E, float *out,
                                                        float *da0, float *da1, int64_t C)
{
    const int64_t groups = (Y + CHUNK_GROUP - 1) / CHUNK_GROUP;
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= groups * E) return;
    const int64_t g = t / E, e = t - g * E;
    const int64_t y1 = min(Y, (g + 1) * CHUNK_GROUP);
    float s = 0.f;
    for (int64_t y = g * CHUNK_GROUP; y < y1; ++y) s += in[y * E + e];
    if (groups > 1) out[g * E + e] = s;
    else if (e < C) da0[e] = s;
    else da1[e - C] = s;
}

void node_bwd_grid(int64_t n, int64_t C, int64_t &gx, int64_t &gy)
{
    gx = (C + 255) / 256;   // columns of a workgroup: 64 lanes x 4 (one lane's worth for V = 1)
    gy = std::max<int64_t>(1, std::min<int64_t>((n + 63) / 64, std::max<int64_t>(1, 2048 / gx)));
}

__global__ void fill_f32_kernel(float *p, int64_t n, float v)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

int wl_of(int wide)
{
    switch (wide) { case 4: return 2; case 8: return 3; case 16: return 4; case 32: return 5; default: return -1; }
}

bool is_float_type(int t) { return t == ATTGNN_F32 || t == ATTGNN_F16 || t == ATTGNN_BF16; }

}  // namespace

extern "C" int attgnn_sddmm_block(const int32_t *rp, const int32_t *col, const int32_t *map,
                                  const int32_t *window_of_part, int64_t parts, const void *A,
                                  int64_t a_head_stride, int64_t a_row_stride, const void *B,
                                  int64_t b_head_stride, int64_t b_row_stride, int in_type,
                                  int64_t heads, int64_t n_rows, int64_t dim, void *out, int out_type,
                                  int64_t nnz, int window, int wide, int mode, attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (parts < 0 || heads <= 0 || n_rows < 0 || dim <= 0 || nnz < 0) return ATTGNN_E_ARG;
    if (!(window == 8 || window == 16)) return ATTGNN_E_ARG;
    const int wl = wl_of(wide);
    if (wl < 0 || !is_float_type(in_type) || !is_float_type(out_type)) return ATTGNN_E_ARG;
    if (parts == 0) return ATTGNN_OK;
    if (!rp || !col || !map || !window_of_part || !A || !B || !out) return ATTGNN_E_ARG;
    int kind;
    if (in_type == ATTGNN_F16) kind = 0;
    else if (in_type == ATTGNN_BF16) kind = 1;
    else kind = mode == ATTGNN_MODE_F32 ? 3 : 2;
    const int eb = type_bytes(in_type);
    const bool vec = dim % 8 == 0 && a_row_stride % 8 == 0 && b_row_stride % 8 == 0 && a_head_stride % 8 == 0 &&
                     b_head_stride % 8 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(B) % 16 == 0 && (eb == 2 || eb == 4);
    SddmmArgs a{rp, col, map, window_of_part, static_cast<const uint8_t *>(A), static_cast<const uint8_t *>(B),
                static_cast<uint8_t *>(out), parts, nnz, n_rows, dim, a_head_stride, a_row_stride,
                b_head_stride, b_row_stride, int(heads), window, wl, in_type, out_type};
    const dim3 grid(unsigned((parts * heads + 3) / 4)), block(256);
#define ATTGNN_SDDMM_CASE(K)                                                                         \
    case K:                                                                                          \
        if (vec) hipLaunchKernelGGL((sddmm_block_kernel<K, true>), grid, block, 0, s, a);            \
        else hipLaunchKernelGGL((sddmm_block_kernel<K, false>), grid, block, 0, s, a);               \
        break;
    switch (kind) {
        ATTGNN_SDDMM_CASE(0)
        ATTGNN_SDDMM_CASE(1)
        ATTGNN_SDDMM_CASE(2)
        ATTGNN_SDDMM_CASE(3)
    }
#undef ATTGNN_SDDMM_CASE
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_node_scores(const void *X, int x_type, int64_t x_head_stride, int64_t x_row_stride,
                                  const void *vec, int vec_type, int64_t heads, int64_t n_rows, int64_t dim,
                                  float *scores, int64_t s_head_stride, int64_t s_row_stride,
                                  attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (heads <= 0 || n_rows < 0 || dim <= 0 || !is_float_type(x_type) || !is_float_type(vec_type))
        return ATTGNN_E_ARG;
    if (n_rows == 0) return ATTGNN_OK;
    if (!X || !vec || !scores) return ATTGNN_E_ARG;
    const uint8_t *xb = static_cast<const uint8_t *>(X), *vb = static_cast<const uint8_t *>(vec);
    const int eb = type_bytes(x_type), ce = 16 / eb;
    const bool aligned16 = (reinterpret_cast<uintptr_t>(X) & 15) == 0 &&
                           (vec_type != x_type || (reinterpret_cast<uintptr_t>(vec) & 15) == 0);
    const bool vec_ok = aligned16 && dim % ce == 0 && x_head_stride % ce == 0 && x_row_stride % ce == 0 &&
                        (heads == 1 || heads == 2 || heads == 4 || heads == 8);
    if (vec_ok) {
        const dim3 grid(unsigned((n_rows * 16 + 255) / 256)), block(256);
#define ATTGNN_NS_CASE(EBV, HV)                                                                                \
        if (eb == EBV && heads == HV)                                                                          \
            hipLaunchKernelGGL((node_score_vec_kernel<EBV, HV, 1>), grid, block, 0, s, xb, x_type, x_head_stride, \
                               x_row_stride, vb, vec_type, n_rows, dim, scores, s_head_stride, s_row_stride,  \
                               nullptr, nullptr);
        ATTGNN_NS_CASE(2, 1) ATTGNN_NS_CASE(2, 2) ATTGNN_NS_CASE(2, 4) ATTGNN_NS_CASE(2, 8)
        ATTGNN_NS_CASE(4, 1) ATTGNN_NS_CASE(4, 2) ATTGNN_NS_CASE(4, 4) ATTGNN_NS_CASE(4, 8)
#undef ATTGNN_NS_CASE
    } else {
        const int64_t threads = heads * n_rows * 16;
        hipLaunchKernelGGL(node_score_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s, xb, x_type,
                           x_head_stride, x_row_stride, vb, vec_type, heads, n_rows, dim, scores, s_head_stride,
                           s_row_stride);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_node_scores_pair(const void *X, int x_type, int64_t x_head_stride, int64_t x_row_stride,
                                       const void *vec0, const void *vec1, int vec_type, int64_t heads,
                                       int64_t n_rows, int64_t dim, float *scores0, float *scores1,
                                       int64_t s_head_stride, int64_t s_row_stride, attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (heads <= 0 || n_rows < 0 || dim <= 0 || !is_float_type(x_type) || !is_float_type(vec_type))
        return ATTGNN_E_ARG;
    if (n_rows == 0) return ATTGNN_OK;
    if (!X || !vec0 || !vec1 || !scores0 || !scores1) return ATTGNN_E_ARG;
    const uint8_t *xb = static_cast<const uint8_t *>(X);
    const uint8_t *v0 = static_cast<const uint8_t *>(vec0), *v1 = static_cast<const uint8_t *>(vec1);
    const int eb = type_bytes(x_type), ce = 16 / eb;
    const bool aligned16 = (reinterpret_cast<uintptr_t>(X) & 15) == 0 &&
                           (vec_type != x_type || ((reinterpret_cast<uintptr_t>(vec0) |
                                                    reinterpret_cast<uintptr_t>(vec1)) & 15) == 0);
    const bool vec_ok = aligned16 && dim % ce == 0 && x_head_stride % ce == 0 && x_row_stride % ce == 0 &&
                        (heads == 1 || heads == 2 || heads == 4 || heads == 8);
    if (!vec_ok) {
        // the generic kernel, once per vector
        int rc = attgnn_node_scores(X, x_type, x_head_stride, x_row_stride, vec0, vec_type, heads, n_rows, dim,
                                    scores0, s_head_stride, s_row_stride, stream);
        if (rc != ATTGNN_OK) return rc;
        return attgnn_node_scores(X, x_type, x_head_stride, x_row_stride, vec1, vec_type, heads, n_rows, dim,
                                  scores1, s_head_stride, s_row_stride, stream);
    }
    const dim3 grid(unsigned((n_rows * 16 + 255) / 256)), block(256);
#define ATTGNN_NS2_CASE(EBV, HV)                                                                                \
    if (eb == EBV && heads == HV)                                                                              \
        hipLaunchKernelGGL((node_score_vec_kernel<EBV, HV, 2>), grid, block, 0, s, xb, x_type, x_head_stride,  \
                           x_row_stride, v0, vec_type, n_rows, dim, scores0, s_head_stride, s_row_stride, v1,  \
                           scores1);
    ATTGNN_NS2_CASE(2, 1) ATTGNN_NS2_CASE(2, 2) ATTGNN_NS2_CASE(2, 4) ATTGNN_NS2_CASE(2, 8)
    ATTGNN_NS2_CASE(4, 1) ATTGNN_NS2_CASE(4, 2) ATTGNN_NS2_CASE(4, 4) ATTGNN_NS2_CASE(4, 8)
#undef ATTGNN_NS2_CASE
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_edge_from_nodes(const int32_t *rp, const int32_t *col, const int32_t *map,
                                      const int32_t *window_of_part, int64_t parts, const float *s0,
                                      int64_t s0_head_stride, int64_t s0_row_stride, const float *s1,
                                      int64_t s1_head_stride, int64_t s1_row_stride, const void *g0,
                                      int g0_type, int64_t heads, int64_t n_rows, void *out, int out_type,
                                      int64_t nnz, int window, int wide, attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (parts < 0 || heads <= 0 || nnz < 0 || n_rows < 0) return ATTGNN_E_ARG;
    if (!(window == 8 || window == 16)) return ATTGNN_E_ARG;
    const int wl = wl_of(wide);
    if (wl < 0 || !is_float_type(out_type)) return ATTGNN_E_ARG;
    if (g0 && !is_float_type(g0_type)) return ATTGNN_E_ARG;
    if (parts == 0) return ATTGNN_OK;
    if (!rp || !col || !map || !window_of_part || !out) return ATTGNN_E_ARG;
    EdgeArgs a{rp, col, map, window_of_part, s0, s1, static_cast<const uint8_t *>(g0), g0_type,
               s0_head_stride, s0_row_stride, s1_head_stride, s1_row_stride, static_cast<uint8_t *>(out),
               parts, nnz, n_rows, int(heads), window, wl, out_type};
    hipLaunchKernelGGL(edge_from_nodes_kernel, dim3(unsigned(parts)), dim3(256), 0, s, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_rowsum_edge_map(const int32_t *rp, const int32_t *map, const int32_t *window_of_part,
                                      int64_t parts, const void *edge_val, int val_type, int64_t nnz,
                                      int64_t heads, float *out, int out_layout, int64_t m_rows, int window,
                                      int wide, attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (parts < 0 || heads <= 0 || nnz < 0 || m_rows < 0) return ATTGNN_E_ARG;
    if (!(window == 8 || window == 16)) return ATTGNN_E_ARG;
    if (!(out_layout == ATTGNN_LAYOUT_HMD || out_layout == ATTGNN_LAYOUT_MHD)) return ATTGNN_E_ARG;
    const int wl = wl_of(wide);
    if (wl < 0 || !is_float_type(val_type)) return ATTGNN_E_ARG;
    if (!out) return ATTGNN_E_ARG;
    const int64_t n = heads * m_rows;
    if (n > 0) {
        hipLaunchKernelGGL(fill_f32_kernel, dim3(unsigned(std::min<int64_t>((n + 255) / 256, 65536))), dim3(256), 0, s,
                           out, n, 0.f);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return int(e);
    }
    if (parts == 0) return ATTGNN_OK;
    if (!rp || !map || !window_of_part || !edge_val) return ATTGNN_E_ARG;
    const int64_t hs = out_layout == ATTGNN_LAYOUT_HMD ? m_rows : 1;
    const int64_t rs = out_layout == ATTGNN_LAYOUT_HMD ? 1 : heads;
    hipLaunchKernelGGL(rowsum_map_kernel, dim3(unsigned(parts)), dim3(256), 0, s, rp, map, window_of_part, parts,
                       static_cast<const uint8_t *>(edge_val), val_type, nnz, int(heads), window, wl, out, hs, rs,
                       m_rows);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_sddmm_edges(const int32_t *rp, const int32_t *col, const int64_t *block_base,
                                  const int32_t *edge_ids, const uint16_t *positions,
                                  const int32_t *window_of_part, int64_t parts, const void *A,
                                  int64_t a_head_stride, int64_t a_row_stride, const void *B,
                                  int64_t b_head_stride, int64_t b_row_stride, int in_type, int64_t heads,
                                  int64_t n_rows, int64_t dim, void *out, int out_type, int64_t nnz, int window,
                                  int mode, attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (parts < 0 || heads <= 0 || n_rows < 0 || dim <= 0 || nnz < 0) return ATTGNN_E_ARG;
    if (!(window == 8 || window == 16)) return ATTGNN_E_ARG;
    if (!is_float_type(in_type) || !is_float_type(out_type)) return ATTGNN_E_ARG;
    if (parts == 0) return ATTGNN_OK;
    if (!rp || !col || !block_base || !edge_ids || !positions || !window_of_part || !A || !B || !out)
        return ATTGNN_E_ARG;
    int kind;
    if (in_type == ATTGNN_F16) kind = 0;
    else if (in_type == ATTGNN_BF16) kind = 1;
    else kind = mode == ATTGNN_MODE_F32 ? 3 : 2;
    const int eb = type_bytes(in_type);
    const bool vec = dim % 8 == 0 && a_row_stride % 8 == 0 && b_row_stride % 8 == 0 && a_head_stride % 8 == 0 &&
                     b_head_stride % 8 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(B) % 16 == 0 && (eb == 2 || eb == 4);
    SddmmArgs a{rp, col, nullptr, window_of_part, static_cast<const uint8_t *>(A), static_cast<const uint8_t *>(B),
                static_cast<uint8_t *>(out), parts, nnz, n_rows, dim, a_head_stride, a_row_stride,
                b_head_stride, b_row_stride, int(heads), window, 5, in_type, out_type};
    const dim3 grid{static_cast<unsigned>((parts + 3) / 4)}, block{256};
    if (vec && eb == 2 && dim <= 128) {
        const int ndc = int((dim + 31) / 32);
#define ATTGNN_BATCHED(K, N) hipLaunchKernelGGL((sddmm_edges_batched_kernel<K, N>), grid, block, 0, s, a, block_base, \
                                                edge_ids, positions)
        if (ndc == 1) { if (kind == 0) ATTGNN_BATCHED(0, 1); else ATTGNN_BATCHED(1, 1); }
        else if (ndc == 2) { if (kind == 0) ATTGNN_BATCHED(0, 2); else ATTGNN_BATCHED(1, 2); }
        else { if (kind == 0) ATTGNN_BATCHED(0, 4); else ATTGNN_BATCHED(1, 4); }
#undef ATTGNN_BATCHED
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? ATTGNN_OK : int(e);
    }
#define ATTGNN_EDGES_CASE(K)                                                                               \
    case K:                                                                                                \
        if (vec) hipLaunchKernelGGL((sddmm_edges_kernel<K, true>), grid, block, 0, s, a, block_base,       \
                                    edge_ids, positions);                                                  \
        else hipLaunchKernelGGL((sddmm_edges_kernel<K, false>), grid, block, 0, s, a, block_base, edge_ids, \
                                positions);                                                                \
        break;
    switch (kind) {
        ATTGNN_EDGES_CASE(0)
        ATTGNN_EDGES_CASE(1)
        ATTGNN_EDGES_CASE(2)
        ATTGNN_EDGES_CASE(3)
    }
#undef ATTGNN_EDGES_CASE
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int64_t attgnn_attention_node_bwd_workspace_elems(int64_t n_rows, int64_t heads, int64_t dim)
{
    if (n_rows < 0 || heads <= 0 || dim <= 0) return -1;
    int64_t gx, gy;
    node_bwd_grid(n_rows, heads * dim, gx, gy);
    // chunk partials, then the partials of groups of CHUNK_GROUP chunks
    return 2 * heads * dim * (gy + (gy + CHUNK_GROUP - 1) / CHUNK_GROUP);
}

extern "C" int attgnn_attention_node_bwd(const void *xs, const void *xt, int x_type, int64_t n_rows, int64_t heads,
                                         int64_t dim, const float *a0, const float *a1, const float *ds0,
                                         const float *ds1, const float *base, void *dxs, void *dxt, float *da0,
                                         float *da1, float *workspace, int64_t workspace_elems,
                                         attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (n_rows < 0 || heads <= 0 || dim <= 0 || !is_float_type(x_type)) return ATTGNN_E_ARG;
    if (!a0 || !a1 || !da0 || !da1 || !workspace) return ATTGNN_E_ARG;
    if (n_rows > 0 && (!xs || !xt || !ds0 || !ds1 || !dxt)) return ATTGNN_E_ARG;
    if (!dxs && xs != xt) return ATTGNN_E_ARG;   // combined output needs one shared input
    if (workspace_elems < attgnn_attention_node_bwd_workspace_elems(n_rows, heads, dim)) return ATTGNN_E_ARG;
    const int64_t C = heads * dim;
    int64_t gx, gy;
    node_bwd_grid(n_rows, C, gx, gy);
    NodeBwdArgs a{static_cast<const uint8_t *>(xs), static_cast<const uint8_t *>(xt), a0, a1, ds0, ds1, base,
                  static_cast<uint8_t *>(dxs), static_cast<uint8_t *>(dxt), workspace, n_rows, heads, dim, C,
                  x_type};
    const int eb = type_bytes(x_type);
    auto al = [&](const void *p, int bytes) { return p == nullptr || reinterpret_cast<uintptr_t>(p) % bytes == 0; };
    const bool vec = dim % 4 == 0 && al(xs, 4 * eb) && al(xt, 4 * eb) && al(dxs, 4 * eb) && al(dxt, 4 * eb) &&
                     al(base, 16);
    if (vec)
        hipLaunchKernelGGL(attn_node_bwd_kernel<4>, dim3(unsigned((C + 255) / 256), unsigned(gy)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(attn_node_bwd_kernel<1>, dim3(unsigned((C + 63) / 64), unsigned(gy)), dim3(256), 0, s, a);
    // two fixed-order levels (gy <= 2048 <= CHUNK_GROUP^2)
    const int64_t E = 2 * C;
    float *level2 = workspace + gy * E;
    int64_t Y = gy;
    const float *src = workspace;
    while (true) {
        const int64_t groups = (Y + CHUNK_GROUP - 1) / CHUNK_GROUP;
        hipLaunchKernelGGL(chunk_sum_kernel, dim3(unsigned((groups * E + 255) / 256)), dim3(256), 0, s, src, Y, E,
                           level2, da0, da1, C);
        if (groups == 1) break;
        src = level2;
        Y = groups;
        level2 = workspace;   // the chunk partials are consumed; reuse their space
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}
