// Window-block SpMM for gfx950: Y = A . X with A stored as condensed window
// tiles (see attgnn.h, attgnn_spmm_block).
//
// Reference counterpart: spmm_forward_cuda_kernel_{fp16,tf32}_balance and their
// gnn launchers (gcn/SpMM/src/spmmKernel.cu:158-265, 959-1075, 1693-1765),
// which run one 8-row part per thread block, feed mma.sync from per-lane
// 4-byte X gathers and combine split windows with float atomicAdd.
//
// This design:
//  * one 64-lane wave per (part, feature slice); a slice is 16*T features,
//    lane n of each 16-lane group owning features [n*T, n*T+T) of the slice;
//  * per k-block of 32 condensed columns each lane gathers the T-element
//    vectors of 8 X rows (one 16/8/4-byte load per row, coalesced across the
//    16 lanes that share a row) and transposes them in registers into T MFMA
//    B operands, one per N tile -- no LDS round trip;
//  * the lane's A operand (8 tile values of window row n) is one vector load
//    from the dense tile, shared by the T MFMAs;
//  * FP32 accumulation in every mode; split windows write partial tiles that a
//    second kernel sums in part order, so results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <mutex>

#include "attgnn.h"
#include "attn.hpp"
#include "mfma.hpp"

using namespace attgnn;

namespace {

enum Kind { K_F16, K_BF16, K_FP8, K_BF16X3, K_F32 };

template <int K> struct KindTraits;
template <> struct KindTraits<K_F16> { static constexpr int EB = 2; };
template <> struct KindTraits<K_BF16> { static constexpr int EB = 2; };
template <> struct KindTraits<K_FP8> { static constexpr int EB = 1; };
template <> struct KindTraits<K_BF16X3> { static constexpr int EB = 4; };
template <> struct KindTraits<K_F32> { static constexpr int EB = 4; };

template <int BYTES>
__device__ __forceinline__ void load_bytes(const uint8_t *p, uint32_t *d)
{
    if constexpr (BYTES == 32) {
        const u32x4 a = reinterpret_cast<const u32x4 *>(p)[0];
        const u32x4 b = reinterpret_cast<const u32x4 *>(p)[1];
        d[0] = a[0]; d[1] = a[1]; d[2] = a[2]; d[3] = a[3];
        d[4] = b[0]; d[5] = b[1]; d[6] = b[2]; d[7] = b[3];
    } else if constexpr (BYTES == 16) {
        const u32x4 a = *reinterpret_cast<const u32x4 *>(p);
        d[0] = a[0]; d[1] = a[1]; d[2] = a[2]; d[3] = a[3];
    } else if constexpr (BYTES == 8) {
        const u32x2 a = *reinterpret_cast<const u32x2 *>(p);
        d[0] = a[0]; d[1] = a[1];
    } else if constexpr (BYTES == 4) {
        d[0] = *reinterpret_cast<const uint32_t *>(p);
    } else if constexpr (BYTES == 2) {
        d[0] = *reinterpret_cast<const uint16_t *>(p);
    } else {
        d[0] = *p;
    }
}

struct SpmmArgs {
    const int32_t *rp, *col, *win, *atomic;
    const uint8_t *val;     // tiles, or per-edge values [heads][val_head_stride] (mapped)
    const int32_t *map;     // mapped variant: CSR position per tile slot, -1 = none; compact map: packed ids
    const uint16_t *masks;  // compact variant: row mask per condensed column
    const int64_t *bbase;   // compact variant: value offset per 32-column block
    const uint8_t *x;       // X rows: element (h, c, f) at x[h*x_head_stride + c*x_row_stride + f]
    void *y;                // Y element (h, r, f) at y[h*y_head_stride + r*y_row_stride + f]
    float *partial;         // fp32 partial tiles of split windows
    const float *row_scale, *col_scale;
    int64_t parts, m_ori, dim_n;
    int64_t x_head_stride, x_row_stride, y_head_stride, y_row_stride, val_head_stride;
    int slices, heads, window, wide_log2, y_type, pipe;
    int xcd_slices;         // 1: slice-per-XCD task mapping (slices divides 8)
    int32_t *split_list;    // [0] = count, then (first, last) part of each split window
    // attention variant: tile value = exp(LeakyReLU(r1[row] + c1[col]) - r2[row] - c2[col])
    // with dropout on the edge id; element (h, v) at [h*hs + v*ns]
    const float *r1, *r2, *c1, *c2;   // r2 / c2 may be null (0)
    int64_t lg_hs, lg_ns, of_hs, of_ns, n_nodes;
    float slope, keep_scale;
    uint32_t thresh;
    uint64_t seed;
    // A_CMAP_DOT: dot_out[h*val_head_stride + e] = <dot_x[h, window row], X[h, column]>
    // for every edge e in the structure; dot_x has X's type and layout and
    // dot_rows rows: element (h, r, f) at dot_x[h*dot_head_stride + r*dot_row_stride + f]
    const uint8_t *dot_x;
    float *dot_out;
    int64_t dot_rows, dot_head_stride, dot_row_stride;
    // striped node order (stripe_pos): with stripe_log2 > 0 window row r is
    // stored at stripe_pos(r); with owners > 1 as well, the tasks are split by
    // the owner of their columns (see the kernel) and oseg[part][o] holds
    // the first column of owner o's segment of each part (o = 0..owners)
    int stripe_log2, owners, groups;
    int64_t stripe_q, x_bytes;
    const int32_t *oseg;
};

// Striped node order with 2^L owners of range q (a multiple of 64): node j
// sits at stripe_pos(j) = ((j >> 6) mod 2^L) q + (j >> (6 + L)) 64 + (j mod 64),
// i.e. 64-node stripes dealt round-robin to the owners, each owner's stripes
// contiguous.  A bijection on [0, 2^L q); 16-row windows stay contiguous.
__host__ __device__ __forceinline__ int64_t stripe_pos(int64_t j, int lg, int64_t q)
{
    return ((j >> 6) & ((1 << lg) - 1)) * q + ((j >> (6 + lg)) << 6) + (j & 63);
}

// Store n consecutive fp32 values as y_type (F32 / F16 / BF16).
template <int N>
__device__ __forceinline__ void store_row(void *base, int64_t idx, int y_type, const float (&v)[N])
{
    if (y_type == ATTGNN_F32) {
        float *d = static_cast<float *>(base) + idx;
        if constexpr (N % 4 == 0) {
#pragma unroll
            for (int j = 0; j < N; j += 4) *reinterpret_cast<f32x4 *>(d + j) = f32x4{v[j], v[j + 1], v[j + 2], v[j + 3]};
        } else {
#pragma unroll
            for (int j = 0; j < N; ++j) d[j] = v[j];
        }
    } else {
        uint16_t h[N];
#pragma unroll
        for (int j = 0; j < N; ++j)
            h[j] = y_type == ATTGNN_F16 ? __builtin_bit_cast(uint16_t, _Float16(v[j])) : uint16_t(bf16_bits_rne(v[j]));
        uint16_t *d = static_cast<uint16_t *>(base) + idx;
        if constexpr (N % 8 == 0) {
#pragma unroll
            for (int j = 0; j < N; j += 8)
                *reinterpret_cast<u32x4 *>(d + j) = u32x4{h[j] | uint32_t(h[j + 1]) << 16, h[j + 2] | uint32_t(h[j + 3]) << 16,
                                                          h[j + 4] | uint32_t(h[j + 5]) << 16, h[j + 6] | uint32_t(h[j + 7]) << 16};
        } else if constexpr (N % 4 == 0) {
#pragma unroll
            for (int j = 0; j < N; j += 4)
                *reinterpret_cast<u32x2 *>(d + j) = u32x2{h[j] | uint32_t(h[j + 1]) << 16, h[j + 2] | uint32_t(h[j + 3]) << 16};
        } else {
#pragma unroll
            for (int j = 0; j < N; ++j) d[j] = h[j];
        }
    }
}

template <int BYTES>
__device__ __forceinline__ void load_bytes_nt(const uint8_t *p, uint32_t *d)
{
    if constexpr (BYTES == 32) {
        const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
        const u32x4 b = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p) + 1);
        d[0] = a[0]; d[1] = a[1]; d[2] = a[2]; d[3] = a[3];
        d[4] = b[0]; d[5] = b[1]; d[6] = b[2]; d[7] = b[3];
    } else if constexpr (BYTES == 16) {
        const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
        d[0] = a[0]; d[1] = a[1]; d[2] = a[2]; d[3] = a[3];
    } else if constexpr (BYTES == 8) {
        const u32x2 a = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(p));
        d[0] = a[0]; d[1] = a[1];
    } else {
        load_bytes<BYTES>(p, d);
    }
}

// A operand: dense value tiles, edge-id map tiles (values gathered by edge
// id), compact value tiles, map tiles with the attention probabilities
// rebuilt, compact edge-id lists (values gathered by edge id), and compact
// edge-id lists with each slot's dot product of the window row of a second
// matrix and the gathered row written per edge (A_CMAP_DOT).
enum AMode { A_DENSE = 0, A_MAPPED = 1, A_COMPACT = 2, A_ATTN = 3, A_CMAP = 4, A_CMAP_DOT = 5 };

// Per-wave state and the three steps of one k-block (32 condensed columns):
// column ids of the lane's 8 k slots, the A operand (tile row n), the X rows.
// A_CMAP_DOT staging in LDS, per wave: the k-block's 32 gathered rows (16T
// 16-bit features each; row c at c * dot_pitch + 16 * ((c >> 3) & 3) bytes, so
// the four 8-column groups of a lane position write to different banks), then
// the block's 16 x 32 edge ids (-1: no edge), rows of 32 ints.
template <int T> constexpr int dot_pitch() { return 32 * T + 16; }
template <int T> constexpr int dot_rows_bytes() { return 32 * dot_pitch<T>() + 48; }
// the id tile's rows are 36 words apart: row n's 16-byte writes by the lanes
// of one 8-column group, and the column reads of one MFMA output row, then
// fall in distinct banks (a 32-word pitch put every row's writes in the same
// 4 banks, an 8-way conflict)
constexpr int ID_PITCH = 36;
template <int T> constexpr int dot_wave_bytes() { return dot_rows_bytes<T>() + 16 * ID_PITCH * 4; }

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int KIND, int T, int AMODE>
struct KBlock {
    static constexpr int EB = KindTraits<KIND>::EB;
    static constexpr int AW = (8 * EB) / 4;
    static constexpr int XB = T * EB;
    static constexpr int XW = XB >= 4 ? XB / 4 : 1;
    static constexpr bool CMP = AMODE == A_COMPACT || AMODE == A_CMAP || AMODE == A_CMAP_DOT;   // mask-addressed
    static constexpr bool IDS = AMODE == A_CMAP || AMODE == A_CMAP_DOT;   // packed entries are edge ids

    const SpmmArgs &a;
    const uint8_t *xh, *vh;
    int c_begin, nreal, plen, g, n, W, wide, head;
    int64_t f0;
    bool f_ok, a_row_ok;
    float r1v, r2v;   // attention variant: the lane's window-row terms
    int *dot_ids = nullptr;  // A_CMAP_DOT: the wave's 16 x 32 edge-id tile in LDS (row pitch ID_PITCH)

    // Column ids: read once per k-block, so nontemporal -- except for the
    // compact 8-features-per-lane instance (width 256, two slices), where the
    // cached load measured 1.3% faster (4.757 -> 4.695 ms on the 114.6M-entry
    // graph) while at 4 features per lane it was 3.5% slower, and for the
    // compact edge lists, whose heads run in one workgroup and share the
    // ids (and row masks) through the CU's L1 (fused attention backward
    // 0.497 -> 0.476 ms at 4 heads x 64, 1.009 -> 0.947 at 8).
    __device__ __forceinline__ void cols(int kb, int (&cid)[8]) const
    {
        const int kg = kb + 8 * g;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr ((AMODE == A_COMPACT && T == 8) || AMODE == A_CMAP || AMODE == A_CMAP_DOT) cid[i] = kg + i < nreal ? a.col[c_begin + kg + i] : -1;
            else cid[i] = kg + i < nreal ? __builtin_nontemporal_load(a.col + c_begin + kg + i) : -1;
        }
    }

    __device__ __forceinline__ void tile(int kb, uint32_t (&av)[AW]) const
    {
        const int kg = kb + 8 * g;
#pragma unroll
        for (int w = 0; w < AW; ++w) av[w] = 0;
        if (!a_row_ok || kg >= nreal) return;
        if constexpr (CMP) {
            return;  // built by tile_compact, which needs the whole wave
        } else if constexpr (AMODE == A_MAPPED) {
            // the 8 edge ids of row n (one or two vector loads), then the 8
            // value gathers issued together
            int m[8];
            if (wide >= 8) {
                const int64_t cc = int64_t(c_begin) + kg;
                const int32_t *mp = a.map + (cc >> a.wide_log2) * W * wide + int64_t(n) * wide + (cc & (wide - 1));
                const u32x4 m0 = *reinterpret_cast<const u32x4 *>(mp);
                const u32x4 m1 = *reinterpret_cast<const u32x4 *>(mp + 4);
#pragma unroll
                for (int i = 0; i < 4; ++i) { m[i] = int(m0[i]); m[i + 4] = int(m1[i]); }
            } else {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int64_t cc = int64_t(c_begin) + kg + 4 * h;
                    u32x4 mv = {~0u, ~0u, ~0u, ~0u};
                    if (kg + 4 * h < plen)
                        mv = *reinterpret_cast<const u32x4 *>(a.map + (cc >> a.wide_log2) * W * wide + int64_t(n) * wide);
#pragma unroll
                    for (int i = 0; i < 4; ++i) m[4 * h + i] = int(mv[i]);
                }
            }
            uint32_t e[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                e[i] = 0;
                if (m[i] >= 0) load_bytes<EB>(vh + int64_t(m[i]) * EB, &e[i]);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if constexpr (EB == 4) av[i] = e[i];
                else if constexpr (EB == 2) av[i >> 1] |= e[i] << (16 * (i & 1));
                else av[i >> 2] |= e[i] << (8 * (i & 3));
            }
        } else if (wide >= 8) {
            const int64_t cc = int64_t(c_begin) + kg;
            const int64_t off = (cc >> a.wide_log2) * W * wide + int64_t(n) * wide + (cc & (wide - 1));
            load_bytes_nt<8 * EB>(a.val + off * EB, av);
        } else {  // wide == 4: two 4-column blocks
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (kg + 4 * h < plen) {
                    const int64_t cc = int64_t(c_begin) + kg + 4 * h;
                    const int64_t off = (cc >> a.wide_log2) * W * wide + int64_t(n) * wide;
                    load_bytes_nt<4 * EB>(a.val + off * EB, av + h * (AW / 2));
                }
            }
        }
    }

    // The lane's entries of row n in the 8 columns (bit n of m[i]) sit at
    // consecutive packed positions from idx: values, or (A_CMAP) edge ids
    // whose values are then gathered, all ids loaded before the gathers.
    __device__ __forceinline__ void packed(const uint32_t (&m)[8], int64_t idx, uint32_t (&av)[AW]) const
    {
        uint32_t e[8];
        if constexpr (IDS) {
            int id[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                id[i] = -1;
                if ((m[i] >> n) & 1u) {
                    id[i] = a.map[idx];
                    ++idx;
                }
            }
            if constexpr (AMODE == A_CMAP_DOT) {
                u32x4 *d = reinterpret_cast<u32x4 *>(dot_ids + n * ID_PITCH + 8 * g);
                d[0] = u32x4{uint32_t(id[0]), uint32_t(id[1]), uint32_t(id[2]), uint32_t(id[3])};
                d[1] = u32x4{uint32_t(id[4]), uint32_t(id[5]), uint32_t(id[6]), uint32_t(id[7])};
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                e[i] = 0;
                if (id[i] >= 0) load_bytes<EB>(vh + int64_t(id[i]) * EB, &e[i]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                e[i] = 0;
                if ((m[i] >> n) & 1u) {
                    load_bytes<EB>(a.val + idx * EB, &e[i]);
                    ++idx;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr (EB == 4) av[i] = e[i];
            else if constexpr (EB == 2) av[i >> 1] |= e[i] << (16 * (i & 1));
            else av[i >> 2] |= e[i] << (8 * (i & 3));
        }
    }

    // Compact tiles: every lane of the wave must call this (shuffles).
    __device__ __forceinline__ void tile_compact(int kb, uint32_t (&av)[AW]) const
    {
        const int kg = kb + 8 * g;
#pragma unroll
        for (int w = 0; w < AW; ++w) av[w] = 0;
        uint32_t m[8];
        int tot = 0, below = 0;
        if (kg < plen) {
            const int64_t cc = int64_t(c_begin) + kg;
            const u32x4 mv = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(a.masks + cc));
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                m[i] = (mv[i >> 1] >> (16 * (i & 1))) & 0xffffu;
                tot += __builtin_popcount(m[i]);
                below += __builtin_popcount(m[i] & ((1u << n) - 1u));
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) m[i] = 0;
        }
        // entries of the earlier 8-column groups of this block
        const int t0 = __shfl(tot, 0), t1 = __shfl(tot, 16), t2 = __shfl(tot, 32);
        const int prefix = (g > 0 ? t0 : 0) + (g > 1 ? t1 : 0) + (g > 2 ? t2 : 0);
        if (kg >= nreal) {
            if constexpr (AMODE == A_CMAP_DOT) {   // no entries: the id tile row stays empty
                u32x4 *d = reinterpret_cast<u32x4 *>(dot_ids + n * ID_PITCH + 8 * g);
                d[0] = u32x4{~0u, ~0u, ~0u, ~0u};
                d[1] = u32x4{~0u, ~0u, ~0u, ~0u};
            }
            return;
        }
        int64_t idx = a.bbase[(int64_t(c_begin) + kb) >> 5] + prefix + below;
        packed(m, idx, av);
    }

    // Compact tiles, prefetch form: the lane's 8 row masks of k-block kb and
    // the block's value offset, loaded one k-block ahead so the value loads of
    // a block depend only on registers.
    __device__ __forceinline__ void masks(int kb, u32x4 &mv, int64_t &base) const
    {
        const int kg = kb + 8 * g;
        mv = u32x4{0u, 0u, 0u, 0u};
        base = 0;
        if (kb < nreal) base = a.bbase[(int64_t(c_begin) + kb) >> 5];
        if (kg < plen) {
            if constexpr (IDS) mv = *reinterpret_cast<const u32x4 *>(a.masks + int64_t(c_begin) + kg);
            else mv = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(a.masks + int64_t(c_begin) + kg));
        }
    }

    // tile_compact from prefetched masks (every lane of the wave must call it).
    __device__ __forceinline__ void tile_compact_pf(int kb, const u32x4 &mv, int64_t base, uint32_t (&av)[AW]) const
    {
        const int kg = kb + 8 * g;
#pragma unroll
        for (int w = 0; w < AW; ++w) av[w] = 0;
        uint32_t m[8];
        int tot = 0, below = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            m[i] = (mv[i >> 1] >> (16 * (i & 1))) & 0xffffu;
            tot += __builtin_popcount(m[i]);
            below += __builtin_popcount(m[i] & ((1u << n) - 1u));
        }
        const int t0 = __shfl(tot, 0), t1 = __shfl(tot, 16), t2 = __shfl(tot, 32);
        const int prefix = (g > 0 ? t0 : 0) + (g > 1 ? t1 : 0) + (g > 2 ? t2 : 0);
        if (kg >= nreal) {
            if constexpr (AMODE == A_CMAP_DOT) {   // no entries: the id tile row stays empty
                u32x4 *d = reinterpret_cast<u32x4 *>(dot_ids + n * ID_PITCH + 8 * g);
                d[0] = u32x4{~0u, ~0u, ~0u, ~0u};
                d[1] = u32x4{~0u, ~0u, ~0u, ~0u};
            }
            return;
        }
        int64_t idx = base + prefix + below;
        packed(m, idx, av);
    }

    // Attention variant: the A operand is rebuilt from node terms where it is
    // used -- exp(LeakyReLU(r1[row] + c1[col]) - r2[row] - c2[col]), dropout
    // keyed on the edge id from the map -- with the edge softmax's own
    // expression and rounding, so it equals the stored probabilities bit for bit.
    __device__ __forceinline__ void tile_attn(int kb, const int (&cid)[8], uint32_t (&av)[AW]) const
    {
        const int kg = kb + 8 * g;
#pragma unroll
        for (int w = 0; w < AW; ++w) av[w] = 0;
        if (!a_row_ok || kg >= nreal) return;
        int m[8];
        if (wide >= 8) {
            const int64_t cc = int64_t(c_begin) + kg;
            const int32_t *mp = a.map + (cc >> a.wide_log2) * W * wide + int64_t(n) * wide + (cc & (wide - 1));
            const u32x4 m0 = *reinterpret_cast<const u32x4 *>(mp);
            const u32x4 m1 = *reinterpret_cast<const u32x4 *>(mp + 4);
#pragma unroll
            for (int i = 0; i < 4; ++i) { m[i] = int(m0[i]); m[i + 4] = int(m1[i]); }
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int64_t cc = int64_t(c_begin) + kg + 4 * h;
                u32x4 mv = {~0u, ~0u, ~0u, ~0u};
                if (kg + 4 * h < plen)
                    mv = *reinterpret_cast<const u32x4 *>(a.map + (cc >> a.wide_log2) * W * wide + int64_t(n) * wide);
#pragma unroll
                for (int i = 0; i < 4; ++i) m[4 * h + i] = int(mv[i]);
            }
        }
        float c1v[8], c2v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const bool ok = m[i] >= 0 && cid[i] >= 0 && cid[i] < a.n_nodes;
            c1v[i] = ok ? a.c1[int64_t(head) * a.lg_hs + int64_t(cid[i]) * a.lg_ns] : 0.f;
            c2v[i] = (ok && a.c2) ? a.c2[int64_t(head) * a.of_hs + int64_t(cid[i]) * a.of_ns] : 0.f;
        }
        uint32_t e[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            e[i] = 0;
            if (m[i] >= 0 && cid[i] >= 0 && cid[i] < a.n_nodes) {
                const float p = attn_prob(r1v + c1v[i], r2v + c2v[i], a.slope, a.thresh, a.keep_scale, a.seed, head,
                                          m[i]);
                if constexpr (KIND == K_F16) e[i] = __builtin_bit_cast(uint16_t, _Float16(p));
                else if constexpr (KIND == K_BF16) e[i] = uint16_t(bf16_bits_rne(p));
                else e[i] = __float_as_uint(p);
            }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr (EB == 4) av[i] = e[i];
            else if constexpr (EB == 2) av[i >> 1] |= e[i] << (16 * (i & 1));
            else av[i >> 2] |= e[i] << (8 * (i & 3));
        }
    }

    __device__ __forceinline__ void load_a(int kb, const int (&cid)[8], uint32_t (&av)[AW]) const
    {
        if constexpr (CMP) tile_compact(kb, av);
        else if constexpr (AMODE == A_ATTN) tile_attn(kb, cid, av);
        else tile(kb, av);
    }

    __device__ __forceinline__ void rows(const int (&cid)[8], uint32_t (&xv)[8][XW]) const
    {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
#pragma unroll
            for (int w = 0; w < XW; ++w) xv[i][w] = 0;
            if (cid[i] >= 0 && f_ok) load_bytes<XB>(xh + (int64_t(cid[i]) * a.x_row_stride + f0) * EB, xv[i]);
        }
    }

    // A_CMAP_DOT: S[r][c] = <xa row r, gathered row c> for the k-block's 32
    // columns, K = the 16T features in chunks of 32 (chunk s2, features
    // 32 s2 + 8g.. at lane group g, as the SDDMM assigns them), stored to
    // dot_out from the MFMA output layout through the wave's id tile (filled by
    // packed() for this block).  The gathered rows are transposed through
    // LDS: the aggregation holds 8 columns x T features per lane, the product
    // needs 8 features of one column.
    __device__ __forceinline__ void dot_block(const uint32_t (&xv)[8][XW], const u32x4 *xa, uint8_t *dg) const
    {
        if constexpr (AMODE == A_CMAP_DOT && T >= 2 && (KIND == K_F16 || KIND == K_BF16)) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int c = 8 * g + i;
                uint32_t *d = reinterpret_cast<uint32_t *>(dg + c * dot_pitch<T>() + 16 * ((c >> 3) & 3) + n * T * 2);
#pragma unroll
                for (int w = 0; w < XW; ++w) d[w] = xv[i][w];
            }
            wave_sync();
            f32x4 sacc[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int c = 16 * t + n;
                const uint8_t *row = dg + c * dot_pitch<T>() + 16 * ((c >> 3) & 3);
                sacc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s2 = 0; s2 < T / 2; ++s2) {
                    const u32x4 B = *reinterpret_cast<const u32x4 *>(row + (32 * s2 + 8 * g) * 2);
                    sacc[t] = KIND == K_F16 ? mfma_f16(xa[s2], B, sacc[t]) : mfma_bf16(xa[s2], B, sacc[t]);
                }
            }
            // product (4g + q, 16t + n) of this lane goes to the edge in that slot
            float *out = a.dot_out + int64_t(head) * a.val_head_stride;
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int e = dot_ids[(4 * g + q) * ID_PITCH + 16 * t + n];
                    if (e >= 0) out[e] = sacc[t][q];
                }
            wave_sync();   // id tile and gathered rows read before the next block writes them
        }
    }

    __device__ __forceinline__ static void mma(const uint32_t (&av)[AW], const uint32_t (&xv)[8][XW], f32x4 (&acc)[T])
    {
        if constexpr (KIND == K_F16 || KIND == K_BF16) {
            const u32x4 A = {av[0], av[1], av[2], av[3]};
#pragma unroll
            for (int j = 0; j < T; ++j) {
                const u32x4 B = gather16<T>(xv, j);
                acc[j] = KIND == K_F16 ? mfma_f16(A, B, acc[j]) : mfma_bf16(A, B, acc[j]);
            }
        } else if constexpr (KIND == K_FP8) {
            const u32x2 A = {av[0], av[1]};
#pragma unroll
            for (int j = 0; j < T; ++j) acc[j] = mfma_fp8(A, gather8<T>(xv, j), acc[j]);
        } else if constexpr (KIND == K_F32) {
#pragma unroll
            for (int j = 0; j < T; ++j)
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    acc[j] = mfma_f32(__uint_as_float(av[i]), __uint_as_float(xv[i][j]), acc[j]);
        } else {  // K_BF16X3
            uint32_t ah[8], al[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) bf16_split(__uint_as_float(av[i]), ah[i], al[i]);
            const u32x4 AH = {ah[0] | ah[1] << 16, ah[2] | ah[3] << 16, ah[4] | ah[5] << 16, ah[6] | ah[7] << 16};
            const u32x4 AL = {al[0] | al[1] << 16, al[2] | al[3] << 16, al[4] | al[5] << 16, al[6] | al[7] << 16};
#pragma unroll
            for (int j = 0; j < T; ++j) {
                uint32_t bh[8], bl[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) bf16_split(__uint_as_float(xv[i][j]), bh[i], bl[i]);
                const u32x4 BH = {bh[0] | bh[1] << 16, bh[2] | bh[3] << 16, bh[4] | bh[5] << 16, bh[6] | bh[7] << 16};
                const u32x4 BL = {bl[0] | bl[1] << 16, bl[2] | bl[3] << 16, bl[4] | bl[5] << 16, bl[6] | bl[7] << 16};
                acc[j] = mfma_bf16(AL, BH, acc[j]);
                acc[j] = mfma_bf16(AH, BL, acc[j]);
                acc[j] = mfma_bf16(AH, BH, acc[j]);
            }
        }
    }
};

// One wave per (part, head, slice).  256-thread workgroups hold 4 tasks.
// MAPPED: tile slot values are val[head][map[slot]] (edge-map SpMM).
// PIPE 0: each k-block loads its column ids, then its rows, then multiplies.
// PIPE 1: the next k-block's column ids load while this block multiplies.
// PIPE 2: two k-blocks per iteration, the next pair's column ids prefetched.
// The 64-feature A_CMAP_DOT instance is compiled for 5 waves per SIMD (86
// VGPRs, no spills; 4 otherwise); every other instance as the compiler picks.
template <int KIND, int T, int AMODE, int PIPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((AMODE == A_CMAP_DOT && T == 4) ? 5 : 1)))
void spmm_block_kernel(SpmmArgs a)
{
    using KB = KBlock<KIND, T, AMODE>;
    constexpr int EB = KB::EB, AW = KB::AW, XW = KB::XW;

    const int lane = threadIdx.x & 63;
    // the wave's index in the workgroup, in a scalar register: the task
    // (part, head, slice) and everything derived from it stay wave-uniform
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int64_t part;
    int head, slice;
    if (a.xcd_slices) {
        // Workgroups are dispatched round-robin over the 8 XCDs.  Give every
        // XCD a single feature slice (slices divides 8) so each XCD's L2 caches
        // one slice of the gathered rows: xcd = b % 8 -> slice xcd % S; the
        // other bits of b and the wave pick the (part, head).  The heads of a
        // part stay in one workgroup: they share column ids and edge ids
        // through the CU's L1 (spreading heads over XCDs measured slower).
        const int64_t b = blockIdx.x;
        const int xcd = int(b & 7), S = a.slices, groups = 8 / S;
        slice = xcd % S;
        const int64_t ph = ((b >> 3) * groups + xcd / S) * 4 + wave;
        if (ph >= a.parts * a.heads) return;
        part = ph / a.heads;
        head = int(ph - part * a.heads);
    } else {
        const int64_t task = int64_t(blockIdx.x) * 4 + wave;
        if (task >= a.parts * a.heads * a.slices) return;
        part = task / (int64_t(a.heads) * a.slices);
        const int hs = int(task - part * a.heads * a.slices);
        head = hs / a.slices;
        slice = hs - head * a.slices;
    }
    // a negative window marks a padding part of a fixed-capacity structure
    if (a.win[part] < 0) return;

    // the wave of a split window's first part (head 0, slice 0) lists the
    // window for the combine; list order varies, each window's sum does not
    if (a.split_list && head == 0 && slice == 0 && lane == 0 && a.atomic[part]) {
        const int w = a.win[part];
        if (!(part > 0 && a.atomic[part - 1] && a.win[part - 1] == w)) {
            int64_t last = part;
            while (last + 1 < a.parts && a.atomic[last + 1] && a.win[last + 1] == w) ++last;
            const int k = atomicAdd(a.split_list, 1);
            a.split_list[1 + 2 * k] = int32_t(part);
            a.split_list[2 + 2 * k] = int32_t(last);
        }
    }

    KB k{a};
    k.xh = a.x + int64_t(head) * a.x_head_stride * EB;
    k.vh = a.val + int64_t(head) * a.val_head_stride * EB;
    k.c_begin = a.rp[2 * part];
    k.nreal = a.rp[2 * part + 1] - k.c_begin;
    k.plen = a.rp[2 * part + 2] - k.c_begin;
    k.g = lane >> 4;
    k.n = lane & 15;
    k.W = a.window;
    k.wide = 1 << a.wide_log2;
    k.head = head;
    k.f0 = int64_t(slice) * 16 * T + int64_t(k.n) * T;
    k.f_ok = k.f0 < a.dim_n;
    k.a_row_ok = k.n < k.W;
    k.r1v = 0.f;
    k.r2v = 0.f;
    if constexpr (AMODE == A_ATTN) {
        const int64_t row = int64_t(a.win[part]) * k.W + k.n;
        if (k.a_row_ok && row < a.n_nodes) {
            k.r1v = a.r1[int64_t(head) * a.lg_hs + row * a.lg_ns];
            if (a.r2) k.r2v = a.r2[int64_t(head) * a.of_hs + row * a.of_ns];
        }
    }
    const int g = k.g, W = k.W, nreal = k.nreal;
    const int64_t f0 = k.f0;

    // A_CMAP_DOT: the window rows of dot_x (MFMA A operand: row n, features
    // 32 s2 + 8g..) for the whole part, and the wave's LDS staging
    constexpr bool DOT = AMODE == A_CMAP_DOT && T >= 2 && (KIND == K_F16 || KIND == K_BF16);
    extern __shared__ __align__(16) uint8_t spmm_lds[];
    u32x4 xa[T >= 2 ? T / 2 : 1];
    uint8_t *dg = nullptr;
    if constexpr (DOT) {
        dg = spmm_lds + wave * dot_wave_bytes<T>();
        k.dot_ids = reinterpret_cast<int *>(dg + dot_rows_bytes<T>());
        const int64_t drow = int64_t(a.win[part]) * W + k.n;
        const bool ok = k.a_row_ok && drow < a.dot_rows;
        const uint8_t *xr = a.dot_x + (int64_t(head) * a.dot_head_stride + drow * a.dot_row_stride) * 2;
#pragma unroll
        for (int s2 = 0; s2 < T / 2; ++s2)
            xa[s2] = ok ? *reinterpret_cast<const u32x4 *>(xr + (32 * s2 + 8 * g) * 2) : u32x4{0u, 0u, 0u, 0u};
    }

    f32x4 acc[T];
#pragma unroll
    for (int j = 0; j < T; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};

    if constexpr (PIPE == 0) {
        for (int kb = 0; kb < nreal; kb += 32) {
            int cid[8];
            uint32_t av[AW], xv[8][XW];
            k.cols(kb, cid);
            k.load_a(kb, cid, av);
            k.rows(cid, xv);
            KB::mma(av, xv, acc);
        }
    } else if constexpr (PIPE == 1 && KB::CMP) {
        // column ids, row masks and value offset of the next k-block load
        // while this one multiplies
        int cid[8];
        u32x4 mv;
        int64_t mb;
        k.cols(0, cid);
        k.masks(0, mv, mb);
        for (int kb = 0; kb < nreal; kb += 32) {
            uint32_t av[AW], xv[8][XW];
            int nid[8];
            u32x4 nmv;
            int64_t nmb;
            k.tile_compact_pf(kb, mv, mb, av);
            k.rows(cid, xv);
            k.cols(kb + 32, nid);
            k.masks(kb + 32, nmv, nmb);
            KB::mma(av, xv, acc);
            if constexpr (DOT) k.dot_block(xv, xa, dg);
#pragma unroll
            for (int i = 0; i < 8; ++i) cid[i] = nid[i];
            mv = nmv;
            mb = nmb;
        }
    } else if constexpr (PIPE == 1) {
        int cid[8];
        k.cols(0, cid);
        for (int kb = 0; kb < nreal; kb += 32) {
            uint32_t av[AW], xv[8][XW];
            int nid[8];
            k.load_a(kb, cid, av);
            k.rows(cid, xv);
            k.cols(kb + 32, nid);
            KB::mma(av, xv, acc);
#pragma unroll
            for (int i = 0; i < 8; ++i) cid[i] = nid[i];
        }
    } else if constexpr (KB::CMP) {
        // two k-blocks per iteration; the next pair's column ids, row masks
        // and value offsets load while this pair multiplies
        int c0[8], c1[8];
        u32x4 m0, m1;
        int64_t b0, b1;
        k.cols(0, c0);
        k.cols(32, c1);
        k.masks(0, m0, b0);
        k.masks(32, m1, b1);
        for (int kb = 0; kb < nreal; kb += 64) {
            uint32_t av0[AW], av1[AW], xv0[8][XW], xv1[8][XW];
            int n0[8], n1[8];
            u32x4 nm0, nm1;
            int64_t nb0, nb1;
            k.tile_compact_pf(kb, m0, b0, av0);
            k.tile_compact_pf(kb + 32, m1, b1, av1);
            k.rows(c0, xv0);
            k.rows(c1, xv1);
            k.cols(kb + 64, n0);
            k.cols(kb + 96, n1);
            k.masks(kb + 64, nm0, nb0);
            k.masks(kb + 96, nm1, nb1);
            KB::mma(av0, xv0, acc);
            KB::mma(av1, xv1, acc);
#pragma unroll
            for (int i = 0; i < 8; ++i) { c0[i] = n0[i]; c1[i] = n1[i]; }
            m0 = nm0; m1 = nm1; b0 = nb0; b1 = nb1;
        }
    } else {
        int c0[8], c1[8];
        k.cols(0, c0);
        k.cols(32, c1);
        for (int kb = 0; kb < nreal; kb += 64) {
            uint32_t av0[AW], av1[AW], xv0[8][XW], xv1[8][XW];
            int n0[8], n1[8];
            k.load_a(kb, c0, av0);
            k.load_a(kb + 32, c1, av1);
            k.rows(c0, xv0);
            k.rows(c1, xv1);
            k.cols(kb + 64, n0);
            k.cols(kb + 96, n1);
            KB::mma(av0, xv0, acc);
            KB::mma(av1, xv1, acc);
#pragma unroll
            for (int i = 0; i < 8; ++i) { c0[i] = n0[i]; c1[i] = n1[i]; }
        }
    }

    // ---- store rows 4g+q of the window, features f0..f0+T-1
    if (!k.f_ok) return;
    const int64_t wrow = int64_t(a.win[part]) * W;
    const bool split = a.atomic[part] != 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = 4 * g + q;
        if (r >= W) break;
        const int64_t row = wrow + r;
        float v[T];
#pragma unroll
        for (int j = 0; j < T; ++j) v[j] = acc[j][q];
        if constexpr (KIND == K_FP8) {
            const float rs = row < a.m_ori ? a.row_scale[row] : 0.f;
#pragma unroll
            for (int j = 0; j < T; ++j) v[j] *= rs * a.col_scale[f0 + j];
        }
        if (split) {
            float *dst = a.partial + ((part * a.heads + head) * W + r) * a.dim_n + f0;
            if constexpr (T % 4 == 0) {
#pragma unroll
                for (int j = 0; j < T; j += 4) *reinterpret_cast<f32x4 *>(dst + j) = f32x4{v[j], v[j + 1], v[j + 2], v[j + 3]};
            } else {
#pragma unroll
                for (int j = 0; j < T; ++j) dst[j] = v[j];
            }
        } else {
            if (row >= a.m_ori) break;
            const int64_t orow = a.stripe_log2 ? stripe_pos(row, a.stripe_log2, a.stripe_q) : row;
            store_row<T>(a.y, int64_t(head) * a.y_head_stride + orow * a.y_row_stride + f0, a.y_type, v);
        }
    }
}

// ---------------------------------------------------------------------------
// Owner-split SpMM over compact F16 / BF16 tiles in striped node order.  Every
// part is taken by K = groups * slices waves (K divides 8), one per XCD slot:
// the wave on XCD x takes feature slice s (T features per lane) of the
// columns owned by the owners of group gi -- gi, gi + groups, ... -- with
// (gi, s) from x mod K (the host picks groups, see
// attgnn_spmm_compact_striped).  Striped column
// ids make each owner's columns one id range, so a part's columns fall into
// consecutive owner segments (oseg) and each XCD's L2 caches its owners' rows
// of one slice (with 4 owners a quarter of X, at 256 FP16 features and at 64;
// half at 256 with slices alone).  Every task writes its fp32 tile;
// spmm_owner_combine_kernel sums them.
//
// The loads go through buffer resources, whose out-of-range reads return
// zero: a column outside the owner's segment (or past the part) gets an
// out-of-range offset instead of a branch, so the k-loop has no divergent
// control flow and no zero-initialised load targets.  Requires X to span
// fewer than 2^31 bytes and rows of fewer than 2^24 bytes (32-bit offsets,
// 24-bit multiplies) and dim_n to be a multiple of 16 T (every lane's
// features inside the row).

constexpr int BUF_WORD3 = 0x00020000;     // raw buffer resource, untyped loads (gfx9 family)
constexpr uint32_t BUF_OOB = 0x80000000u; // past every buffer the owner kernel reads: the load returns 0

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, int(bytes), BUF_WORD3);
}

// A k-block's structure, loaded one block ahead and spread over the wave
// once (not once per lane): lane 16g + j (j < 8) holds the id of column
// kb + 8g + j, lane 16g + j (j < 4) the row masks of columns kb + 8g + 2j and
// kb + 8g + 2j + 1; base is the block's first packed value.
struct OwnPre {
    uint32_t id, mk;
    int64_t base;
};

// The value of lane 16g + I for every lane of group g (ds_swizzle bit mode,
// and-mask 0x10, or-mask I: within each 32-lane half the source keeps the
// lane's group bit).
template <int I>
__device__ __forceinline__ uint32_t from_group_lane(uint32_t v)
{
    return uint32_t(__builtin_amdgcn_ds_swizzle(int(v), 0x10 | (I << 5)));
}

template <int KIND, int T>
__global__ __launch_bounds__(256) void spmm_owner_kernel(SpmmArgs a)
{
    using KB = KBlock<KIND, T, A_COMPACT>;
    constexpr int AW = KB::AW, XW = KB::XW;
    static_assert((KIND == K_F16 || KIND == K_BF16) && (T == 4 || T == 8), "16-bit tiles, 4 or 8 features per lane");

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // K = groups * slices tasks per part on XCD slots x mod K (with K < 8 the
    // XCDs x and x + K hold the same task kind and take alternate parts);
    // slot (group gi, slice s) takes slice s of owners gi, gi + groups, ...
    const int xcd = int(blockIdx.x & 7), K = a.groups * a.slices;
    const int slot = xcd % K, slice = slot % a.slices, group = slot / a.slices;
    const int64_t part = ((int64_t(blockIdx.x >> 3) * (8 / K)) + xcd / K) * 4 + wave;
    if (part >= a.parts || a.win[part] < 0) return;
    const int c_begin = a.rp[2 * part], plen = a.rp[2 * part + 2] - c_begin;
    int lo = 0, hi = 0;   // the owner segment being walked, relative to c_begin
    const int g = lane >> 4, n = lane & 15;
    const uint32_t rb = uint32_t(a.x_row_stride) * 2;                  // bytes per X row
    const uint32_t fb = uint32_t(slice * 16 * T + n * T) * 2;          // the lane's features in a row
    const __amdgpu_buffer_rsrc_t rx = buf_rsrc(a.x, uint32_t(a.x_bytes));
    const __amdgpu_buffer_rsrc_t rc = buf_rsrc(a.col + c_begin, uint32_t(plen) * 4);
    const __amdgpu_buffer_rsrc_t rm = buf_rsrc(a.masks + c_begin, uint32_t(plen) * 2);
    const uint32_t id_off = n < 8 ? uint32_t(8 * g + n) * 4 : BUF_OOB;
    const uint32_t mk_off = n < 4 ? uint32_t(8 * g + 2 * n) * 2 : BUF_OOB;

    auto pre = [&](int kb, OwnPre &p) {
        p.id = __builtin_amdgcn_raw_buffer_load_b32(rc, id_off + uint32_t(kb) * 4, 0, 0);
        p.mk = __builtin_amdgcn_raw_buffer_load_b32(rm, mk_off + uint32_t(kb) * 2, 0, 0);
        p.base = kb < plen ? a.bbase[(int64_t(c_begin) + kb) >> 5] : 0;
    };

    // A k-block is issued (its value and X-row loads started), then consumed
    // (A operand formed, MFMAs) after the structure two blocks ahead is
    // requested, so each block waits once, for its own loads.  Only columns
    // in [lo, hi) contribute; FULL: the whole block lies in [lo, hi) (every
    // block but a segment's first and last), so no column needs the range
    // test.  The
    // group's masks and ids come from lanes 16g + w / 16g + i
    // (from_group_lane); the block's packed values are loaded 64 at a time,
    // one per lane, and each lane takes its row's entries from the lanes that
    // loaded them (ds_bpermute).
    struct Blk {
        uint32_t mine, rowbits, v;   // the lane's present columns, all its row's columns, its loaded value
        int start, total;            // first packed entry of the lane's row; entries in the block
        int64_t base;
    };
    const uint32_t low2 = ((1u << n) - 1u) * 0x10001u;   // rows below n, in both halves
    auto issue = [&](auto full, int kb, const OwnPre &p, Blk &b, uint32_t (&xv)[8][XW]) {
        constexpr bool FULL = decltype(full)::value;
        const uint32_t mk[4] = {from_group_lane<0>(p.mk), from_group_lane<1>(p.mk), from_group_lane<2>(p.mk),
                                from_group_lane<3>(p.mk)};
        uint32_t tot = 0, below = 0, rowbits = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            tot += __builtin_popcount(mk[w]);
            below += __builtin_popcount(mk[w] & low2);
            rowbits |= ((mk[w] >> n) & 1u) << (2 * w);
            rowbits |= ((mk[w] >> (16 + n)) & 1u) << (2 * w + 1);
        }
        // packed order of a block: 8-column groups, then rows, then columns
        const int t0 = __shfl(int(tot), 0), t1 = __shfl(int(tot), 16), t2 = __shfl(int(tot), 32);
        const int t3 = __shfl(int(tot), 48);
        b.total = __builtin_amdgcn_readfirstlane(t0 + t1 + t2 + t3);
        b.start = (g > 0 ? t0 : 0) + (g > 1 ? t1 : 0) + (g > 2 ? t2 : 0) + int(below);
        uint32_t live = 0xffu;
        if constexpr (!FULL) {
            const int kg = kb + 8 * g;
            const int l8 = min(max(lo - kg, 0), 8), h8 = min(max(hi - kg, 0), 8);
            live = ((1u << h8) - 1u) & ~((1u << l8) - 1u);
        }
        b.rowbits = rowbits;
        b.mine = rowbits & live;
        b.base = p.base;
        b.v = __builtin_amdgcn_raw_buffer_load_b16(buf_rsrc(a.val + p.base * 2, uint32_t(b.total) * 2),
                                                   uint32_t(lane) * 2, 0, 0);
        // ids below 2^24 (X spans < 2^31 bytes of >= 128-byte rows): 24-bit multiplies
        const uint32_t ids[8] = {from_group_lane<0>(p.id), from_group_lane<1>(p.id), from_group_lane<2>(p.id),
                                 from_group_lane<3>(p.id), from_group_lane<4>(p.id), from_group_lane<5>(p.id),
                                 from_group_lane<6>(p.id), from_group_lane<7>(p.id)};
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t off = FULL || ((live >> i) & 1u) ? __umul24(ids[i], rb) + fb : BUF_OOB;
            if constexpr (T == 8) {
                const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rx, off, 0, 0);
#pragma unroll
                for (int w = 0; w < 4; ++w) xv[i][w] = v[w];
            } else {
                const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rx, off, 0, 0);
                xv[i][0] = v[0];
                xv[i][1] = v[1];
            }
        }
    };
    auto issue_any = [&](int kb, const OwnPre &p, Blk &b, uint32_t (&xv)[8][XW]) {
        if (kb >= lo && kb + 32 <= hi) issue(std::true_type{}, kb, p, b, xv);
        else issue(std::false_type{}, kb, p, b, xv);
    };

    f32x4 acc[T];
#pragma unroll
    for (int j = 0; j < T; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    auto consume = [&](const Blk &b, const uint32_t (&xv)[8][XW]) {
        uint32_t e[8], src[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            src[i] = uint32_t(b.start) + __builtin_popcount(b.rowbits & ((1u << i) - 1u));
            const uint32_t got = uint32_t(__builtin_amdgcn_ds_bpermute(int(src[i] * 4), int(b.v)));
            e[i] = ((b.mine >> i) & 1u) && src[i] < 64u ? got : 0u;
        }
        for (int c = 64; c < b.total; c += 64) {   // uniform; blocks of more than 64 entries only
            const uint32_t v = __builtin_amdgcn_raw_buffer_load_b16(
                buf_rsrc(a.val + b.base * 2, uint32_t(b.total) * 2), uint32_t(c + lane) * 2, 0, 0);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const uint32_t got = uint32_t(__builtin_amdgcn_ds_bpermute(int((src[i] - c) * 4), int(v)));
                if (((b.mine >> i) & 1u) && src[i] - uint32_t(c) < 64u) e[i] = got;
            }
        }
        uint32_t av[AW];
#pragma unroll
        for (int w = 0; w < AW; ++w) av[w] = e[2 * w] | (e[2 * w + 1] << 16);
        KB::mma(av, xv, acc);
    };

    // one block's loads in flight while the structure of the next two is
    // prefetched (p0: block kb, p1: block kb + 32)
    for (int o = group; o < a.owners; o += a.groups) {
        const int64_t so = part * (a.owners + 1) + o;
        lo = a.oseg[so] - c_begin;
        hi = a.oseg[so + 1] - c_begin;
        const int kb0 = lo & ~31;
        OwnPre p0, p1;
        pre(kb0, p0);
        pre(kb0 + 32, p1);
        for (int kb = kb0; kb < hi; kb += 32) {
            Blk b;
            uint32_t xv[8][XW];
            issue_any(kb, p0, b, xv);
            p0 = p1;
            pre(kb + 64, p1);
            consume(b, xv);
        }
    }

    // the fp32 tile as 16-byte records [part][slot][4 rows x T/4][lane]
    f32x4 *dst = reinterpret_cast<f32x4 *>(a.partial) + (part * K + slot) * T * 64 + lane;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < T; j += 4)
            __builtin_nontemporal_store(f32x4{acc[j][q], acc[j + 1][q], acc[j + 2][q], acc[j + 3][q]},
                                        dst + (q * (T / 4) + j / 4) * 64);
}

// Sum of an owner-split SpMM's tiles: one thread per (part, slice, lane); the
// thread of a window's first part adds the tiles of the window's parts in
// part order, owner groups 0..groups-1 within each part, and stores the window's
// rows 4g+q (g, n = lane / 16, lane % 16), features 16 T slice + T n.. at
// their stripe positions.
template <int T>
__global__ __launch_bounds__(256) void spmm_owner_combine_kernel(SpmmArgs a)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int S = a.slices;
    const int64_t p = t / (64 * S);
    if (p >= a.parts) return;
    const int rem = int(t - p * 64 * S);
    const int s = rem >> 6, lane = rem & 63;
    const int w = a.win[p];
    if (w < 0 || (p > 0 && a.win[p - 1] == w)) return;
    int64_t last = p;
    while (last + 1 < a.parts && a.win[last + 1] == w) ++last;
    float v[4][T];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < T; ++j) v[q][j] = 0.f;
    for (int64_t pp = p; pp <= last; ++pp)
        for (int gi = 0; gi < a.groups; ++gi) {
            const f32x4 *src = reinterpret_cast<const f32x4 *>(a.partial) + (pp * a.groups * S + gi * S + s) * T * 64 + lane;
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int j = 0; j < T; j += 4) {
                    const f32x4 r = __builtin_nontemporal_load(src + (q * (T / 4) + j / 4) * 64);
                    v[q][j] += r[0];
                    v[q][j + 1] += r[1];
                    v[q][j + 2] += r[2];
                    v[q][j + 3] += r[3];
                }
        }
    const int g = lane >> 4, n = lane & 15;
    const int64_t f0 = int64_t(s) * 16 * T + int64_t(n) * T;
    if (f0 >= a.dim_n) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = 4 * g + q;
        if (r >= a.window) break;
        const int64_t row = int64_t(w) * a.window + r;
        if (row >= a.m_ori) break;
        store_row<T>(a.y, stripe_pos(row, a.stripe_log2, a.stripe_q) * a.y_row_stride + f0, a.y_type, v[q]);
    }
}

// Sum the partial tiles of the listed split windows in part order, V features
// per thread over every (window, row, feature group): the sum of each output
// element runs q = first..last from 0, whatever the list order.
template <int V>
__global__ __launch_bounds__(256) void spmm_combine_list_kernel(SpmmArgs a)
{
    const int head = blockIdx.y;
    const int64_t n_split = a.split_list[0];
    const int64_t groups = a.dim_n / V, per = int64_t(a.window) * groups;
    const int64_t total = n_split * per;
    for (int64_t it = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; it < total;
         it += int64_t(gridDim.x) * blockDim.x) {
        const int64_t k = it / per, e = it - k * per;
        const int64_t p = a.split_list[1 + 2 * k], last = a.split_list[2 + 2 * k];
        const int64_t w = a.win[p];
        const int64_t r = e / groups, f = (e - r * groups) * V;
        if (w * a.window + r >= a.m_ori) continue;
        float sv[V];
#pragma unroll
        for (int j = 0; j < V; ++j) sv[j] = 0.f;
        for (int64_t q = p; q <= last; ++q) {
            const float *src = a.partial + ((q * a.heads + head) * a.window + r) * a.dim_n + f;
            if constexpr (V == 4) {
                const f32x4 v = *reinterpret_cast<const f32x4 *>(src);
#pragma unroll
                for (int j = 0; j < 4; ++j) sv[j] += v[j];
            } else {
                sv[0] += src[0];
            }
        }
        const int64_t row = w * a.window + r;
        const int64_t orow = a.stripe_log2 ? stripe_pos(row, a.stripe_log2, a.stripe_q) : row;
        store_row<V>(a.y, int64_t(head) * a.y_head_stride + orow * a.y_row_stride + f, a.y_type, sv);
    }
}

__global__ void zero_count_kernel(int32_t *count)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) count[0] = 0;
}

__global__ __launch_bounds__(256) void fill_zero_kernel(u32x4 *p, int64_t n4, uint8_t *tail, int64_t ntail)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t k = i; k < n4; k += stride) p[k] = u32x4{0u, 0u, 0u, 0u};
    for (int64_t k = i; k < ntail; k += stride) tail[k] = 0;
}

int kind_of(int mode, int tile_type, int x_type)
{
    if (mode == ATTGNN_MODE_NATIVE) {
        if (tile_type == ATTGNN_F16 && x_type == ATTGNN_F16) return K_F16;
        if (tile_type == ATTGNN_BF16 && x_type == ATTGNN_BF16) return K_BF16;
        if (tile_type == ATTGNN_FP8E4M3 && x_type == ATTGNN_FP8E4M3) return K_FP8;
        return -1;
    }
    if (tile_type != ATTGNN_F32 || x_type != ATTGNN_F32) return -1;
    if (mode == ATTGNN_MODE_BF16X3) return K_BF16X3;
    if (mode == ATTGNN_MODE_F32) return K_F32;
    return -1;
}

// k-loop schedule from the ATTGNN_SPMM_PIPE_* bits.  Default: two k-blocks per
// iteration for 8/16-bit dense or compact tiles (measured on a 114.6M-entry
// graph: width 64 dense 1.62 -> 1.56 ms, width 256 compact 4.91 -> 4.79 ms);
// one block with prefetched column ids for fp32 data, whose operands take
// twice the registers, and for edge-map tiles (4 heads x 64 on 2.48M entries:
// 0.271 ms with one block, 0.288 ms with two).
// Default k-loop schedule: two blocks per iteration for 8/16-bit value tiles;
// for edge-map tiles only in launches of few waves (one head: the deeper
// prefetch hides more latency per wave), one block otherwise (more waves per
// SIMD fit).  Measured on 2.48M entries x 64 features, compact edge lists:
// 1 head 94 -> 86 us with two blocks, 4 heads 222 -> 254 us.
constexpr int64_t FEW_WAVES = 16384;

int pipe_of(int flags, int eb, bool mapped, int64_t waves = FEW_WAVES)
{
    switch ((flags >> 4) & 3) {
    case 3: return 0;
    case 2: return 2;
    case 1: return 1;
    default: return (eb <= 2 && (!mapped || waves < FEW_WAVES)) ? 2 : 1;
    }
}

// One feature slice per XCD (see the kernel's task mapping) when the slice
// count is above 1 and divides 8, unless the caller turns it off.
int xcd_mapping(int flags, int slices)
{
    return !(flags & ATTGNN_SPMM_NO_XCD_SLICES) && slices > 1 && 8 % slices == 0;
}

// Features per lane: the widest power of two dividing dim_n that keeps a lane
// vector <= 16 bytes, narrowed while half of it still covers dim_n.
int pick_t(int64_t dim_n, int eb)
{
    int t = 1;
    while (t < 8 && t * 2 * eb <= 16 && dim_n % (t * 2) == 0) t *= 2;
    while (t > 1 && int64_t(16) * (t / 2) >= dim_n) t /= 2;
    return t;
}

// Dynamic LDS of an A_CMAP_DOT launch (4 waves per workgroup).
size_t dot_lds_bytes(int t)
{
    switch (t) {
    case 2: return 4 * size_t(dot_wave_bytes<2>());
    case 4: return 4 * size_t(dot_wave_bytes<4>());
    case 8: return 4 * size_t(dot_wave_bytes<8>());
    default: return 0;
    }
}

template <int KIND, int AMODE, int PIPE>
hipError_t launch_pipe(int t, const SpmmArgs &a, hipStream_t s)
{
    const size_t lds = AMODE == A_CMAP_DOT ? dot_lds_bytes(t) : 0;
    const int64_t tasks = a.parts * a.heads * a.slices;
    int64_t blocks = (tasks + 3) / 4;
    if (a.xcd_slices) {
        const int64_t per_round = int64_t(8 / a.slices) * 4;   // (part, head) pairs per 8 workgroups
        blocks = (a.parts * a.heads + per_round - 1) / per_round * 8;
    }
    const dim3 grid{static_cast<unsigned>(blocks)}, block{256};
    switch (t) {
    case 1: hipLaunchKernelGGL((spmm_block_kernel<KIND, 1, AMODE, PIPE>), grid, block, lds, s, a); break;
    case 2: hipLaunchKernelGGL((spmm_block_kernel<KIND, 2, AMODE, PIPE>), grid, block, lds, s, a); break;
    case 4: hipLaunchKernelGGL((spmm_block_kernel<KIND, 4, AMODE, PIPE>), grid, block, lds, s, a); break;
    case 8:
        if constexpr (KindTraits<KIND>::EB <= 2) {
            hipLaunchKernelGGL((spmm_block_kernel<KIND, 8, AMODE, PIPE>), grid, block, lds, s, a);
            break;
        }
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int KIND, int AMODE>
hipError_t launch_main(int t, const SpmmArgs &a, hipStream_t s)
{
    switch (a.pipe) {
    case 0: return launch_pipe<KIND, AMODE, 0>(t, a, s);
    case 2: return launch_pipe<KIND, AMODE, 2>(t, a, s);
    default: return launch_pipe<KIND, AMODE, 1>(t, a, s);
    }
}

int y_bytes(int y_type) { return y_type == ATTGNN_F32 ? 4 : 2; }

// Zero Y (one contiguous buffer of heads * m_ori * dim_n elements of y_type)
// unless the caller guarantees every row is stored.
hipError_t zero_y(const SpmmArgs &a, bool covered, hipStream_t s)
{
    const int64_t nbytes = int64_t(a.heads) * a.m_ori * a.dim_n * y_bytes(a.y_type);
    if (nbytes > 0 && !covered) {
        uint8_t *yb = static_cast<uint8_t *>(a.y);
        const int64_t n4 = reinterpret_cast<uintptr_t>(yb) % 16 == 0 ? nbytes / 16 : 0;
        const int64_t tail = nbytes - 16 * n4;
        const int64_t blocks = std::min<int64_t>((std::max<int64_t>(n4, tail) + 255) / 256, 65536);
        hipLaunchKernelGGL(fill_zero_kernel, dim3(unsigned(blocks)), dim3(256), 0, s,
                           reinterpret_cast<u32x4 *>(yb), n4, yb + 16 * n4, tail);
        return hipGetLastError();
    }
    return hipSuccess;
}

// Zero Y unless covered, run the main kernel, then combine split windows.
hipError_t run_spmm(int kind, int amode, int t, const SpmmArgs &a, bool covered, hipStream_t s)
{
    hipError_t ez = zero_y(a, covered, s);
    if (ez != hipSuccess) return ez;
    if (a.parts == 0) return hipSuccess;
    // the split-window list sits after the partial tiles in the workspace
    SpmmArgs b = a;
    b.split_list = reinterpret_cast<int32_t *>(a.partial + a.parts * a.heads * int64_t(a.window) * a.dim_n);
    hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(64), 0, s, b.split_list);
    hipError_t e;
#define ATTGNN_BY_AMODE(K)                                                                              \
    (amode == A_MAPPED ? launch_main<K, A_MAPPED>(t, b, s)                                          \
                       : amode == A_COMPACT ? launch_main<K, A_COMPACT>(t, b, s)                    \
                       : amode == A_CMAP ? launch_main<K, A_CMAP>(t, b, s)                          \
                                            : amode == A_ATTN ? launch_main<K, A_ATTN>(t, b, s)     \
                                                              : launch_main<K, A_DENSE>(t, b, s))
    switch (kind) {
    case K_F16: e = amode == A_CMAP_DOT ? launch_pipe<K_F16, A_CMAP_DOT, 1>(t, b, s) : ATTGNN_BY_AMODE(K_F16); break;
    case K_BF16: e = amode == A_CMAP_DOT ? launch_pipe<K_BF16, A_CMAP_DOT, 1>(t, b, s) : ATTGNN_BY_AMODE(K_BF16); break;
    case K_FP8: e = amode == A_COMPACT ? launch_main<K_FP8, A_COMPACT>(t, b, s) : launch_main<K_FP8, A_DENSE>(t, b, s); break;
    case K_BF16X3: e = ATTGNN_BY_AMODE(K_BF16X3); break;
    default: e = ATTGNN_BY_AMODE(K_F32); break;
    }
#undef ATTGNN_BY_AMODE
    if (e != hipSuccess) return e;
    // at most parts / 2 split windows; enough workgroups for one element group per thread
    const int v = a.dim_n % 4 == 0 ? 4 : 1;
    const int64_t items = (a.parts / 2 + 1) * int64_t(a.window) * (a.dim_n / v);
    const int64_t cblocks = std::max<int64_t>(1, std::min<int64_t>((items + 255) / 256, 4096));
    if (v == 4)
        hipLaunchKernelGGL(spmm_combine_list_kernel<4>, dim3(unsigned(cblocks), unsigned(a.heads)), dim3(256), 0, s, b);
    else
        hipLaunchKernelGGL(spmm_combine_list_kernel<1>, dim3(unsigned(cblocks), unsigned(a.heads)), dim3(256), 0, s, b);
    return hipGetLastError();
}

template <int KIND, int T>
hipError_t launch_owners(const SpmmArgs &a, hipStream_t s)
{
    const int64_t per_round = int64_t(8 / (a.groups * a.slices)) * 4;   // parts per 8 workgroups
    const dim3 grid{static_cast<unsigned>((a.parts + per_round - 1) / per_round * 8)}, block{256};
    hipLaunchKernelGGL((spmm_owner_kernel<KIND, T>), grid, block, 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t threads = a.parts * a.slices * 64;
    hipLaunchKernelGGL(spmm_owner_combine_kernel<T>, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Owner-split SpMM (compact F16 / BF16 tiles, 4 or 8 features per lane): every
// task writes its tile to the workspace, then the combine stores Y.
hipError_t run_spmm_owners(int kind, int t, const SpmmArgs &a, bool covered, hipStream_t s)
{
    hipError_t e = zero_y(a, covered, s);
    if (e != hipSuccess || a.parts == 0) return e;
    if (kind == K_F16) return t == 8 ? launch_owners<K_F16, 8>(a, s) : launch_owners<K_F16, 4>(a, s);
    return t == 8 ? launch_owners<K_BF16, 8>(a, s) : launch_owners<K_BF16, 4>(a, s);
}

}  // namespace

extern "C" int64_t attgnn_spmm_workspace_elems(int64_t parts, int window, int64_t dim_n)
{
    // partial tiles of split windows, then the list of split windows
    return parts * window * dim_n + parts + 2;
}

namespace {

int wide_log2_of(int wide)
{
    switch (wide) { case 4: return 2; case 8: return 3; case 16: return 4; case 32: return 5; default: return -1; }
}

int elt_bytes(int kind) { return kind == K_FP8 ? 1 : (kind == K_F16 || kind == K_BF16) ? 2 : 4; }

}  // namespace

namespace {

// Argument checks and launch of attgnn_spmm_block (shared with the
// allocation-free variant, which passes its fixed scratch as the workspace).
int spmm_block_run(const int32_t *rp, const int32_t *col, const void *val, int tile_type,
                   const int32_t *window_of_part, const int32_t *atomic, int64_t parts, const void *X, int x_type,
                   int64_t n_x, void *Y, int y_type, int64_t m_ori, int64_t dim_n, int window, int wide, int mode,
                   const float *row_scale, const float *col_scale, float *workspace, int64_t workspace_elems,
                   int flags, attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (parts < 0 || m_ori < 0 || dim_n <= 0 || n_x < 0) return ATTGNN_E_ARG;
    if (!(window == 8 || window == 16)) return ATTGNN_E_ARG;
    const int wl = wide_log2_of(wide);
    if (wl < 0) return ATTGNN_E_ARG;
    const int kind = kind_of(mode, tile_type, x_type);
    if (kind < 0) return ATTGNN_E_UNSUPPORTED;
    if (kind == K_FP8 && (!row_scale || !col_scale)) return ATTGNN_E_ARG;
    if (!Y) return ATTGNN_E_ARG;
    if (!(y_type == ATTGNN_F32 || y_type == ATTGNN_F16 || y_type == ATTGNN_BF16)) return ATTGNN_E_ARG;
    if (parts > 0 && (!rp || !col || !val || !window_of_part || !atomic || !X || !workspace)) return ATTGNN_E_ARG;
    if (parts > 0 && workspace_elems < attgnn_spmm_workspace_elems(parts, window, dim_n)) return ATTGNN_E_ARG;

    const int eb = elt_bytes(kind);
    int t = pick_t(dim_n, eb);
    // requested features per lane (flags bits 8-11): a power of two dividing
    // dim_n, at most the automatic choice
    const int want = (flags >> 8) & 15;
    if (want && want <= t && (want & (want - 1)) == 0 && dim_n % want == 0) t = want;
    if (reinterpret_cast<uintptr_t>(X) % (t * eb) != 0) return ATTGNN_E_ARG;
    if (reinterpret_cast<uintptr_t>(Y) % 16 != 0) return ATTGNN_E_ARG;

    SpmmArgs a{};
    a.rp = rp; a.col = col; a.win = window_of_part; a.atomic = atomic;
    a.val = static_cast<const uint8_t *>(val); a.map = nullptr;
    a.x = static_cast<const uint8_t *>(X);
    a.y = Y; a.partial = workspace;
    a.row_scale = row_scale; a.col_scale = col_scale;
    a.parts = parts; a.m_ori = m_ori; a.dim_n = dim_n;
    a.x_head_stride = 0; a.x_row_stride = dim_n; a.y_head_stride = 0; a.y_row_stride = dim_n;
    a.val_head_stride = 0;
    a.slices = int((dim_n + 16 * t - 1) / (16 * t));
    a.heads = 1; a.window = window; a.wide_log2 = wl; a.y_type = y_type;
    a.pipe = pipe_of(flags, eb, false);
    a.xcd_slices = xcd_mapping(flags, a.slices);
    const hipError_t e = run_spmm(kind, A_DENSE, t, a, (flags & ATTGNN_SPMM_ROWS_COVERED) != 0, s);
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

}  // namespace

extern "C" int attgnn_spmm_block(const int32_t *rp, const int32_t *col, const void *val, int tile_type,
                                 const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                                 const void *X, int x_type, int64_t n_x, void *Y, int y_type, int64_t m_ori,
                                 int64_t dim_n, int window, int wide, int mode, const float *row_scale,
                                 const float *col_scale, float *workspace, int64_t workspace_elems,
                                 int flags, attgnn_stream_t stream)
{
    return spmm_block_run(rp, col, val, tile_type, window_of_part, atomic, parts, X, x_type, n_x, Y, y_type, m_ori,
                          dim_n, window, wide, mode, row_scale, col_scale, workspace, workspace_elems, flags, stream);
}

// ---------------------------------------------------------------------------
// attgnn_spmm_block without a workspace argument, for callers that cannot
// allocate.  Its scratch -- the split windows' partial tiles and the list of
// split windows -- is this fixed array: every device the library runs on holds
// one copy (ATTGNN_NOALLOC_SCRATCH_ELEMS floats, 64 MiB) until the process
// ends.  In host code `scratch` only names the symbol: its device address is
// looked up once per device with hipGetSymbolAddress and that pointer is what
// the kernels receive.  Calls on one device share the array, so calls that
// overlap in time on one device race; the split-window count is cleared by a
// kernel at the start of every call and every partial tile is written before
// it is read, so nothing of an earlier call's contents is used.
static __device__ __align__(256) float scratch[ATTGNN_NOALLOC_SCRATCH_ELEMS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
float *scratch_address[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call there).
hipError_t scratch_on_current_device(float **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_address[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        scratch_address[dev] = static_cast<float *>(p);
    }
    *out = scratch_address[dev];
    return hipSuccess;
}

}  // namespace

extern "C" int attgnn_spmm_block_noalloc(const int32_t *rp, const int32_t *col, const void *val, int tile_type,
                                         const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                                         const void *X, int x_type, int64_t n_x, void *Y, int y_type,
                                         int64_t m_ori, int64_t dim_n, int window, int wide, int mode,
                                         const float *row_scale, const float *col_scale, int flags,
                                         attgnn_stream_t stream)
{
    float *ws = nullptr;
    if (parts > 0 && (window == 8 || window == 16) && dim_n > 0) {
        if (attgnn_spmm_workspace_elems(parts, window, dim_n) > ATTGNN_NOALLOC_SCRATCH_ELEMS)
            return ATTGNN_E_SCRATCH;
        const hipError_t e = scratch_on_current_device(&ws);
        if (e != hipSuccess) return int(e);
    }
    return spmm_block_run(rp, col, val, tile_type, window_of_part, atomic, parts, X, x_type, n_x, Y, y_type, m_ori,
                          dim_n, window, wide, mode, row_scale, col_scale, ws, ATTGNN_NOALLOC_SCRATCH_ELEMS, flags,
                          stream);
}

namespace {

// Edge-value SpMM over an edge-id map (dense map tiles, or compact masks,
// block offsets and packed ids): argument checks and launch.
int edge_map_run(int amode, const int32_t *rp, const int32_t *col, const int32_t *map, const uint16_t *masks,
                 const int64_t *bbase, const void *edge_val, int val_type, int64_t nnz,
                 const int32_t *window_of_part, const int32_t *atomic, int64_t parts, const void *X, int x_type,
                 int x_layout, int64_t heads, int64_t n_x, void *Y, int y_type, int64_t m_ori, int64_t dim_n,
                 int out_layout, int window, int wide, int mode, float *workspace, int64_t workspace_elems,
                 int flags, hipStream_t s, const void *dot_x = nullptr, int64_t dot_rows = 0,
                 float *dot_out = nullptr)
{
    if (parts < 0 || m_ori < 0 || dim_n <= 0 || n_x < 0 || heads <= 0 || nnz < 0) return ATTGNN_E_ARG;
    if (!(window == 8 || window == 16)) return ATTGNN_E_ARG;
    if (!(out_layout == ATTGNN_LAYOUT_HMD || out_layout == ATTGNN_LAYOUT_MHD)) return ATTGNN_E_ARG;
    if (!(x_layout == ATTGNN_LAYOUT_HMD || x_layout == ATTGNN_LAYOUT_MHD)) return ATTGNN_E_ARG;
    const int wl = wide_log2_of(wide);
    if (wl < 0) return ATTGNN_E_ARG;
    const int kind = kind_of(mode, val_type, x_type);
    if (kind < 0 || kind == K_FP8) return ATTGNN_E_UNSUPPORTED;
    if (!Y || !(y_type == ATTGNN_F32 || y_type == ATTGNN_F16 || y_type == ATTGNN_BF16)) return ATTGNN_E_ARG;
    if (parts > 0 && (!rp || !col || !map || !edge_val || !window_of_part || !atomic || !X || !workspace))
        return ATTGNN_E_ARG;
    if ((amode == A_CMAP || amode == A_CMAP_DOT) && parts > 0 && (!masks || !bbase)) return ATTGNN_E_ARG;
    if (parts > 0 && workspace_elems < heads * attgnn_spmm_workspace_elems(parts, window, dim_n)) return ATTGNN_E_ARG;
    const int eb = elt_bytes(kind);
    const int t = pick_t(dim_n, eb);
    if (reinterpret_cast<uintptr_t>(X) % (t * eb) != 0) return ATTGNN_E_ARG;
    if (reinterpret_cast<uintptr_t>(Y) % 16 != 0) return ATTGNN_E_ARG;

    SpmmArgs a{};
    a.rp = rp; a.col = col; a.win = window_of_part; a.atomic = atomic;
    a.val = static_cast<const uint8_t *>(edge_val); a.map = map; a.masks = masks; a.bbase = bbase;
    a.x = static_cast<const uint8_t *>(X);
    a.y = Y; a.partial = workspace;
    a.parts = parts; a.m_ori = m_ori; a.dim_n = dim_n;
    if (x_layout == ATTGNN_LAYOUT_HMD) { a.x_head_stride = n_x * dim_n; a.x_row_stride = dim_n; }
    else { a.x_head_stride = dim_n; a.x_row_stride = heads * dim_n; }
    a.val_head_stride = nnz;
    if (out_layout == ATTGNN_LAYOUT_HMD) { a.y_head_stride = m_ori * dim_n; a.y_row_stride = dim_n; }
    else { a.y_head_stride = dim_n; a.y_row_stride = heads * dim_n; }
    a.slices = int((dim_n + 16 * t - 1) / (16 * t));
    a.heads = int(heads); a.window = window; a.wide_log2 = wl; a.y_type = y_type;
    a.pipe = pipe_of(flags, eb, true, parts * heads * a.slices);
    a.xcd_slices = xcd_mapping(flags, a.slices);
    if (amode == A_CMAP_DOT) {
        // 16-bit types, one feature slice of 32, 64 or 128 features, the
        // one-block schedule; dot_x 16-byte aligned in X's type and layout
        if (!(kind == K_F16 || kind == K_BF16) || !(dim_n == 32 || dim_n == 64 || dim_n == 128) ||
            t != dim_n / 16 || a.slices != 1)
            return ATTGNN_E_UNSUPPORTED;
        if (dot_rows < 0 || (parts > 0 && (!dot_x || !dot_out)) ||
            reinterpret_cast<uintptr_t>(dot_x) % 16 != 0)
            return ATTGNN_E_ARG;
        a.dot_x = static_cast<const uint8_t *>(dot_x);
        a.dot_out = dot_out;
        a.dot_rows = dot_rows;
        // dot_x has X's layout with its own row count: in [H, rows, D] the head
        // stride is dot_rows * D, in [rows, H, D] it is D as for X
        if (x_layout == ATTGNN_LAYOUT_HMD) { a.dot_head_stride = dot_rows * dim_n; a.dot_row_stride = dim_n; }
        else { a.dot_head_stride = dim_n; a.dot_row_stride = heads * dim_n; }
        a.pipe = 1;
        a.xcd_slices = 0;
    }
    const hipError_t e = run_spmm(kind, amode, t, a, (flags & ATTGNN_SPMM_ROWS_COVERED) != 0, s);
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

}  // namespace

extern "C" int attgnn_spmm_edge_map(const int32_t *rp, const int32_t *col, const int32_t *map,
                                    const void *edge_val, int val_type, int64_t nnz,
                                    const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                                    const void *X, int x_type, int x_layout, int64_t heads, int64_t n_x,
                                    void *Y, int y_type, int64_t m_ori, int64_t dim_n, int out_layout, int window,
                                    int wide, int mode, float *workspace, int64_t workspace_elems, int flags,
                                    attgnn_stream_t stream)
{
    return edge_map_run(A_MAPPED, rp, col, map, nullptr, nullptr, edge_val, val_type, nnz, window_of_part, atomic,
                        parts, X, x_type, x_layout, heads, n_x, Y, y_type, m_ori, dim_n, out_layout, window, wide,
                        mode, workspace, workspace_elems, flags, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int attgnn_spmm_edge_map_compact(const int32_t *rp, const int32_t *col, const uint16_t *masks,
                                            const int64_t *block_base, const int32_t *ids, const void *edge_val,
                                            int val_type, int64_t nnz, const int32_t *window_of_part,
                                            const int32_t *atomic, int64_t parts, const void *X, int x_type,
                                            int x_layout, int64_t heads, int64_t n_x, void *Y, int y_type,
                                            int64_t m_ori, int64_t dim_n, int out_layout, int window, int mode,
                                            float *workspace, int64_t workspace_elems, int flags,
                                            attgnn_stream_t stream)
{
    if (window > 16) return ATTGNN_E_ARG;
    return edge_map_run(A_CMAP, rp, col, ids, masks, block_base, edge_val, val_type, nnz, window_of_part, atomic,
                        parts, X, x_type, x_layout, heads, n_x, Y, y_type, m_ori, dim_n, out_layout, window, 32,
                        mode, workspace, workspace_elems, flags, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int attgnn_spmm_edge_map_compact_dot(const int32_t *rp, const int32_t *col, const uint16_t *masks,
                                                const int64_t *block_base, const int32_t *ids, const void *edge_val,
                                                int val_type, int64_t nnz, const int32_t *window_of_part,
                                                const int32_t *atomic, int64_t parts, const void *X, int x_type,
                                                int x_layout, int64_t heads, int64_t n_x, void *Y, int y_type,
                                                int64_t m_ori, int64_t dim_n, int out_layout, int window,
                                                const void *dot_x, int64_t dot_rows, float *dot_out,
                                                float *workspace, int64_t workspace_elems, int flags,
                                                attgnn_stream_t stream)
{
    if (window > 16) return ATTGNN_E_ARG;
    return edge_map_run(A_CMAP_DOT, rp, col, ids, masks, block_base, edge_val, val_type, nnz, window_of_part,
                        atomic, parts, X, x_type, x_layout, heads, n_x, Y, y_type, m_ori, dim_n, out_layout, window,
                        32, ATTGNN_MODE_NATIVE, workspace, workspace_elems, flags,
                        reinterpret_cast<hipStream_t>(stream), dot_x, dot_rows, dot_out);
}

extern "C" int attgnn_spmm_attention(const int32_t *rp, const int32_t *col, const int32_t *map,
                                     const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                                     const attgnn_attn_terms *terms, const void *X, int x_type, int x_layout,
                                     int64_t heads, int64_t n_x, void *Y, int y_type, int64_t m_ori, int64_t dim_n,
                                     int out_layout, int window, int wide, int mode, float *workspace,
                                     int64_t workspace_elems, int flags, attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (parts < 0 || m_ori < 0 || dim_n <= 0 || n_x < 0 || heads <= 0 || !terms) return ATTGNN_E_ARG;
    if (!(window == 8 || window == 16)) return ATTGNN_E_ARG;
    if (!(out_layout == ATTGNN_LAYOUT_HMD || out_layout == ATTGNN_LAYOUT_MHD)) return ATTGNN_E_ARG;
    if (!(x_layout == ATTGNN_LAYOUT_HMD || x_layout == ATTGNN_LAYOUT_MHD)) return ATTGNN_E_ARG;
    if (terms->n_nodes < 0 || !(terms->drop_p >= 0.f && terms->drop_p < 1.f)) return ATTGNN_E_ARG;
    const int wl = wide_log2_of(wide);
    if (wl < 0) return ATTGNN_E_ARG;
    const int kind = kind_of(mode, x_type, x_type);
    if (kind < 0 || kind == K_FP8) return ATTGNN_E_UNSUPPORTED;
    if (!Y || !(y_type == ATTGNN_F32 || y_type == ATTGNN_F16 || y_type == ATTGNN_BF16)) return ATTGNN_E_ARG;
    if (parts > 0 && (!rp || !col || !map || !window_of_part || !atomic || !X || !workspace || !terms->row_logit ||
                      !terms->col_logit))
        return ATTGNN_E_ARG;
    if (parts > 0 && workspace_elems < heads * attgnn_spmm_workspace_elems(parts, window, dim_n)) return ATTGNN_E_ARG;
    const int eb = elt_bytes(kind);
    const int t = pick_t(dim_n, eb);
    if (reinterpret_cast<uintptr_t>(X) % (t * eb) != 0) return ATTGNN_E_ARG;
    if (reinterpret_cast<uintptr_t>(Y) % 16 != 0) return ATTGNN_E_ARG;

    SpmmArgs a{};
    a.rp = rp; a.col = col; a.win = window_of_part; a.atomic = atomic; a.map = map;
    a.x = static_cast<const uint8_t *>(X);
    a.y = Y; a.partial = workspace;
    a.parts = parts; a.m_ori = m_ori; a.dim_n = dim_n;
    if (x_layout == ATTGNN_LAYOUT_HMD) { a.x_head_stride = n_x * dim_n; a.x_row_stride = dim_n; }
    else { a.x_head_stride = dim_n; a.x_row_stride = heads * dim_n; }
    if (out_layout == ATTGNN_LAYOUT_HMD) { a.y_head_stride = m_ori * dim_n; a.y_row_stride = dim_n; }
    else { a.y_head_stride = dim_n; a.y_row_stride = heads * dim_n; }
    a.slices = int((dim_n + 16 * t - 1) / (16 * t));
    a.heads = int(heads); a.window = window; a.wide_log2 = wl; a.y_type = y_type; a.pipe = pipe_of(flags, eb, true);
    a.xcd_slices = xcd_mapping(flags, a.slices);
    a.r1 = terms->row_logit; a.r2 = terms->row_offset; a.c1 = terms->col_logit; a.c2 = terms->col_offset;
    a.lg_hs = terms->logit_head_stride; a.lg_ns = terms->logit_node_stride;
    a.of_hs = terms->offset_head_stride; a.of_ns = terms->offset_node_stride;
    a.n_nodes = terms->n_nodes;
    a.slope = terms->slope;
    a.thresh = terms->drop_p <= 0.f ? 0u : uint32_t(std::min(double(terms->drop_p), 0.999999) * 4294967296.0);
    a.keep_scale = 1.f / (1.f - terms->drop_p);
    a.seed = terms->seed;
    const hipError_t e = run_spmm(kind, A_ATTN, t, a, (flags & ATTGNN_SPMM_ROWS_COVERED) != 0, s);
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_spmm_compact(const int32_t *rp, const int32_t *col, const uint16_t *masks,
                                   const int64_t *block_base, const void *values, int tile_type,
                                   const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                                   const void *X, int x_type, int64_t n_x, void *Y, int y_type, int64_t m_ori,
                                   int64_t dim_n, int window, int mode, const float *row_scale,
                                   const float *col_scale, float *workspace, int64_t workspace_elems, int flags,
                                   attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (parts < 0 || m_ori < 0 || dim_n <= 0 || n_x < 0) return ATTGNN_E_ARG;
    if (!(window == 8 || window == 16)) return ATTGNN_E_ARG;
    const int kind = kind_of(mode, tile_type, x_type);
    if (kind < 0) return ATTGNN_E_UNSUPPORTED;
    if (kind == K_FP8 && (!row_scale || !col_scale)) return ATTGNN_E_ARG;
    if (!Y || !(y_type == ATTGNN_F32 || y_type == ATTGNN_F16 || y_type == ATTGNN_BF16)) return ATTGNN_E_ARG;
    if (parts > 0 && (!rp || !col || !masks || !block_base || !values || !window_of_part || !atomic || !X ||
                      !workspace))
        return ATTGNN_E_ARG;
    if (parts > 0 && workspace_elems < attgnn_spmm_workspace_elems(parts, window, dim_n)) return ATTGNN_E_ARG;
    if (reinterpret_cast<uintptr_t>(masks) % 16 != 0) return ATTGNN_E_ARG;
    const int eb = elt_bytes(kind);
    int t = pick_t(dim_n, eb);
    const int want = (flags >> 8) & 15;
    if (want && want <= t && (want & (want - 1)) == 0 && dim_n % want == 0) t = want;
    if (reinterpret_cast<uintptr_t>(X) % (t * eb) != 0) return ATTGNN_E_ARG;
    if (reinterpret_cast<uintptr_t>(Y) % 16 != 0) return ATTGNN_E_ARG;

    SpmmArgs a{};
    a.rp = rp; a.col = col; a.win = window_of_part; a.atomic = atomic;
    a.val = static_cast<const uint8_t *>(values); a.map = nullptr;
    a.masks = masks; a.bbase = block_base;
    a.x = static_cast<const uint8_t *>(X);
    a.y = Y; a.partial = workspace;
    a.row_scale = row_scale; a.col_scale = col_scale;
    a.parts = parts; a.m_ori = m_ori; a.dim_n = dim_n;
    a.x_head_stride = 0; a.x_row_stride = dim_n; a.y_head_stride = 0; a.y_row_stride = dim_n;
    a.val_head_stride = 0;
    a.slices = int((dim_n + 16 * t - 1) / (16 * t));
    a.heads = 1; a.window = window; a.wide_log2 = 5; a.y_type = y_type;
    a.pipe = pipe_of(flags, eb, false);
    a.xcd_slices = xcd_mapping(flags, a.slices);
    const hipError_t e = run_spmm(kind, A_COMPACT, t, a, (flags & ATTGNN_SPMM_ROWS_COVERED) != 0, s);
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

// ---------------------------------------------------------------------------
// Striped node order (stripe_pos) and the owner-split SpMM.

namespace {

int log2_owners(int owners)
{
    switch (owners) { case 2: return 1; case 4: return 2; case 8: return 3; case 16: return 4; case 32: return 5;
    default: return -1; }
}

// First position of each owner's segment in every part's real columns
// (ascending within a part): a binary search per (part, owner).
__global__ __launch_bounds__(256) void stripe_segments_kernel(const int32_t *rp, const int32_t *win,
                                                              const int32_t *col, int64_t parts, int owners,
                                                              int64_t q, int32_t *seg)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t p = t / (owners + 1);
    if (p >= parts) return;
    const int o = int(t - p * (owners + 1));
    const int c0 = rp[2 * p], c1 = rp[2 * p + 1];
    if (win[p] < 0 || o == 0) { seg[t] = c0; return; }
    if (o == owners) { seg[t] = c1; return; }
    const int64_t target = int64_t(o) * q;
    int lo = c0, hi = c1;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (col[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    seg[t] = lo;
}

constexpr int64_t OWNER_TILE_ELEMS = 8 * 64 * 4 * 8;   // per part: 8 tasks x 64 lanes x 4 rows x T <= 8

}  // namespace

extern "C" int attgnn_stripe_segments_device(const int32_t *rp, const int32_t *window_of_part, const int32_t *col,
                                             int64_t parts, int owners, int64_t m, int32_t *owner_seg,
                                             attgnn_stream_t stream)
{
    const int lg = log2_owners(owners);
    if (lg < 0 || parts < 0 || m <= 0 || m % (int64_t(64) << lg) != 0) return ATTGNN_E_ARG;
    if (parts == 0) return ATTGNN_OK;
    if (!rp || !window_of_part || !col || !owner_seg) return ATTGNN_E_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t threads = parts * (owners + 1);
    hipLaunchKernelGGL(stripe_segments_kernel, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s, rp,
                       window_of_part, col, parts, owners, m >> lg, owner_seg);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int64_t attgnn_spmm_striped_workspace_elems(int64_t parts, int window, int64_t dim_n, int owners)
{
    (void)owners;
    return std::max<int64_t>(attgnn_spmm_workspace_elems(parts, window, dim_n), parts * OWNER_TILE_ELEMS);
}

extern "C" int attgnn_spmm_compact_striped(const int32_t *rp, const int32_t *col, const uint16_t *masks,
                                           const int64_t *block_base, const void *values, int tile_type,
                                           const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                                           const int32_t *owner_seg, int owners, const void *X, int x_type,
                                           int64_t n_x, void *Y, int y_type, int64_t m, int64_t dim_n, int window,
                                           int mode, float *workspace, int64_t workspace_elems, int flags,
                                           attgnn_stream_t stream)
{
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int lg = log2_owners(owners);
    if (lg < 0 || parts < 0 || m <= 0 || dim_n <= 0 || n_x < m) return ATTGNN_E_ARG;
    if (m % (int64_t(64) << lg) != 0) return ATTGNN_E_ARG;
    if (!(window == 8 || window == 16)) return ATTGNN_E_ARG;
    const int kind = kind_of(mode, tile_type, x_type);
    if (kind < 0 || kind == K_FP8) return ATTGNN_E_UNSUPPORTED;
    if (!Y || !(y_type == ATTGNN_F32 || y_type == ATTGNN_F16 || y_type == ATTGNN_BF16)) return ATTGNN_E_ARG;
    if (parts > 0 && (!rp || !col || !masks || !block_base || !values || !window_of_part || !atomic || !X ||
                      !workspace))
        return ATTGNN_E_ARG;
    if (parts > 0 && workspace_elems < attgnn_spmm_striped_workspace_elems(parts, window, dim_n, owners))
        return ATTGNN_E_ARG;
    if (reinterpret_cast<uintptr_t>(masks) % 16 != 0) return ATTGNN_E_ARG;
    const int eb = elt_bytes(kind);
    int t = pick_t(dim_n, eb);
    const int want = (flags >> 8) & 15;
    if (want && want <= t && (want & (want - 1)) == 0 && dim_n % want == 0) t = want;
    if (reinterpret_cast<uintptr_t>(X) % (t * eb) != 0) return ATTGNN_E_ARG;
    if (reinterpret_cast<uintptr_t>(Y) % 16 != 0) return ATTGNN_E_ARG;

    SpmmArgs a{};
    a.rp = rp; a.col = col; a.win = window_of_part; a.atomic = atomic;
    a.val = static_cast<const uint8_t *>(values); a.map = nullptr;
    a.masks = masks; a.bbase = block_base;
    a.x = static_cast<const uint8_t *>(X);
    a.y = Y; a.partial = workspace;
    a.parts = parts; a.m_ori = m; a.dim_n = dim_n;
    a.x_head_stride = 0; a.x_row_stride = dim_n; a.y_head_stride = 0; a.y_row_stride = dim_n;
    a.val_head_stride = 0;
    a.slices = int((dim_n + 16 * t - 1) / (16 * t));
    a.heads = 1; a.window = window; a.wide_log2 = 5; a.y_type = y_type;
    a.pipe = pipe_of(flags, eb, false);
    a.stripe_log2 = lg; a.stripe_q = m >> lg;
    a.x_bytes = n_x * dim_n * eb;
    // tasks per part: slices x owner groups, at most one per XCD.  Automatic
    // group count: the smallest power of two that leaves each XCD about 16 MB
    // of X (its owners' rows of its slice) -- on the 114.6M-entry graph 4
    // groups at 256 FP16 features (15 MB; 3.70 ms against 4.38 with 2 and 5.23
    // with 1) and 2 at 64 (15 MB; 1.18 ms against 1.26 with 4 and 1.40 with
    // 8), with 8 owners.  A requested count (flags bits 12-14) may only be
    // smaller than min(owners, 8 / slices).
    const int max_groups = std::max(1, std::min(owners, 8 / std::max(a.slices, 1)));
    a.groups = 1;
    while (a.groups < max_groups && a.x_bytes > int64_t(a.groups) * a.slices * (int64_t(16) << 20)) a.groups *= 2;
    const int want_groups = (flags >> 12) & 7;
    if (want_groups) a.groups = std::min(max_groups, 1 << (want_groups - 1));
    const bool split = (kind == K_F16 || kind == K_BF16) && (t == 4 || t == 8) && 8 % (a.slices * a.groups) == 0 &&
                       dim_n % (16 * t) == 0 && a.x_bytes < (int64_t(1) << 31) && dim_n * eb < (1 << 24) &&
                       !(flags & ATTGNN_SPMM_NO_XCD_SLICES);
    hipError_t e;
    if (split) {
        if (parts > 0 && !owner_seg) return ATTGNN_E_ARG;
        a.owners = owners;
        a.oseg = owner_seg;
        e = run_spmm_owners(kind, t, a, (flags & ATTGNN_SPMM_ROWS_COVERED) != 0, s);
    } else {
        a.xcd_slices = xcd_mapping(flags, a.slices);
        e = run_spmm(kind, A_COMPACT, t, a, (flags & ATTGNN_SPMM_ROWS_COVERED) != 0, s);
    }
    return e == hipSuccess ? ATTGNN_OK : int(e);
}
