// Window-block preprocessing on the device (gfx950).
//
// Reference counterpart: ATT_Block_gpu.preprocess_gpu_fs (built from
// Block_gpu/DTCSpMM.cpp + DTCSpMM_kernel.cu, Block_gpu/setup.py:10-14; those
// sources are not in the reference tree; caller
// end2end_overhead/gat_final/mgat_csr_ov/mdataset_fp16.py:53-58), which returns
// the structure of blockProcess_sddmm_balance_gnn_trans, and the host windowing
// of gcn/Block_gcn/example.cpp:48-348.  The result is bit-identical to
// attgnn_blocks_build (host) for every source, window, wide, part size and
// empty-window rule, including unsorted rows and duplicate entries.
//
// Pipeline (one stream, no host round trip until the caller needs sizes):
//  plan  1. rows_check   row pointers monotone, row[0] = 0, row[rows] = nnz;
//        2. cols_check   one wave per row: column ids in range, and a flag for
//                        rows whose columns are not ascending;
//        3. window_unique one workgroup per window: its distinct columns in
//                        ascending order and, for every entry, the rank of its
//                        column among them -- a bitonic sort in LDS for small
//                        windows, or a bitmap of the window's column span in LDS
//                        with per-word prefix counts (O(1) rank) for large ones;
//        4. window_scan  one workgroup: prefix sums of padded column counts and
//                        part counts over windows, totals for the caller;
//  fill  5. window_fill  one workgroup per window: condensed columns, tile
//                        initialisation and scatter, parts;
//  compact 6. count (row masks + entry counts per 32-column block) + scan,
//          7. pack (values and positions in block / 8-column group / row order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include <rocprim/device/device_scan.hpp>

#include "attgnn.h"
#include "halfconv.hpp"

namespace {

constexpr int THREADS = 256;
constexpr int SORT_CAP = 4096;   // windows up to this many entries may be sorted in LDS
constexpr int BM_WORDS = 4096;   // bitmap chunk: 131,072 columns per pass
constexpr int ST_FORMAT = 1, ST_UNSUPPORTED = 2;

// Workspace carve-up (256-byte aligned sections).
struct Ws {
    int32_t *status;   // [1]  bit flags ST_*
    int32_t *nuniq;    // [nwin]
    int64_t *ncol;     // [nwin + 1] exclusive prefix of padded column counts
    int64_t *nparts;   // [nwin + 1] exclusive prefix of part counts
    int32_t *rowflag;  // [rows] 1 = columns of the row not ascending
    int32_t *rank;     // [nnz] rank of the entry's column among its window's distinct columns
    int32_t *uniq;     // [nnz] window i's distinct columns at [row[i*W], ...)
};

inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

int64_t ws_bytes(int64_t rows, int64_t nnz, int window)
{
    const int64_t nwin = rows / window;
    return align256(64) + align256(4 * nwin + 4) + 2 * align256(8 * (nwin + 1)) + align256(4 * rows + 4) +
           2 * align256(4 * nnz + 4);
}

Ws ws_carve(void *base, int64_t rows, int64_t nnz, int window)
{
    const int64_t nwin = rows / window;
    uint8_t *p = static_cast<uint8_t *>(base);
    Ws w;
    w.status = reinterpret_cast<int32_t *>(p); p += align256(64);
    w.nuniq = reinterpret_cast<int32_t *>(p); p += align256(4 * nwin + 4);
    w.ncol = reinterpret_cast<int64_t *>(p); p += align256(8 * (nwin + 1));
    w.nparts = reinterpret_cast<int64_t *>(p); p += align256(8 * (nwin + 1));
    w.rowflag = reinterpret_cast<int32_t *>(p); p += align256(4 * rows + 4);
    w.rank = reinterpret_cast<int32_t *>(p); p += align256(4 * nnz + 4);
    w.uniq = reinterpret_cast<int32_t *>(p);
    return w;
}

inline int type_size(int t)
{
    switch (t) {
    case ATTGNN_F32: case ATTGNN_I32: return 4;
    case ATTGNN_F16: case ATTGNN_BF16: return 2;
    default: return 0;
    }
}

inline unsigned grid_for(int64_t n, int64_t per_block = 1, int64_t cap = 1 << 20)
{
    int64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return unsigned(g);
}

// ---- block-wide helpers (256 threads = 4 waves) ---------------------------

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T n = __shfl_up(v, d, 64);
        if (lane >= d) v += n;
    }
    return v;
}

// Exclusive scan of v over the workgroup; `total` receives the sum.  `red`
// holds 4 elements of LDS; the call contains two barriers.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *red, T &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const T incl = wave_incl_scan(v);
    if (lane == 63) red[wave] = incl;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < THREADS / 64; ++w) {
        if (w < wave) off += red[w];
        tot += red[w];
    }
    __syncthreads();
    total = tot;
    return off + incl - v;
}

__device__ __forceinline__ void block_minmax(int &mn, int &mx, int *red)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        mn = min(mn, __shfl_xor(mn, d, 64));
        mx = max(mx, __shfl_xor(mx, d, 64));
    }
    if (lane == 0) { red[2 * wave] = mn; red[2 * wave + 1] = mx; }
    __syncthreads();
    mn = red[0]; mx = red[1];
#pragma unroll
    for (int w = 1; w < THREADS / 64; ++w) { mn = min(mn, red[2 * w]); mx = max(mx, red[2 * w + 1]); }
    __syncthreads();
}

// Blocks per part split (example.cpp:112-152) in closed form: part k of a
// window with `blocks` blocks of `wide` columns, `v` real columns.
__device__ __forceinline__ int64_t part_count_d(int64_t blocks, int part_size, int emit_empty)
{
    if (blocks == 0) return emit_empty ? 1 : 0;
    if (blocks <= part_size) return 1;
    return blocks / part_size + (blocks % part_size ? 1 : 0);
}

// ---- 1. row pointers -------------------------------------------------------

__global__ __launch_bounds__(THREADS) void init_kernel(Ws ws, int64_t rows)
{
    const int64_t t = int64_t(blockIdx.x) * THREADS + threadIdx.x;
    if (t == 0) ws.status[0] = 0;
    for (int64_t j = t; j < rows; j += int64_t(gridDim.x) * THREADS) ws.rowflag[j] = 0;
}

__global__ __launch_bounds__(THREADS) void rows_check_kernel(const int32_t *row, int64_t rows, int64_t nnz, Ws ws)
{
    const int64_t t = int64_t(blockIdx.x) * THREADS + threadIdx.x;
    if (t == 0 && (row[0] != 0 || int64_t(row[rows]) != nnz)) atomicOr(ws.status, ST_FORMAT);
    for (int64_t j = t; j < rows; j += int64_t(gridDim.x) * THREADS)
        if (row[j + 1] < row[j]) atomicOr(ws.status, ST_FORMAT);
}

// ---- 2. column ids and per-row order flags (one wave per row) -------------

__global__ __launch_bounds__(THREADS) void cols_check_kernel(const int32_t *row, const int32_t *col, int64_t rows, Ws ws)
{
    if (ws.status[0]) return;
    const int lane = threadIdx.x & 63;
    const int64_t nw = int64_t(gridDim.x) * (THREADS / 64);
    for (int64_t j = int64_t(blockIdx.x) * (THREADS / 64) + (threadIdx.x >> 6); j < rows; j += nw) {
        const int64_t b = row[j], e = row[j + 1];
        int bad = 0, unsorted = 0;
        for (int64_t m = b + lane; m < e; m += 64) {
            const int32_t c = col[m];
            bad |= (c < 0 || int64_t(c) >= rows);
            if (m > b && col[m - 1] > c) unsorted = 1;
        }
        if (__any(bad) && lane == 0) atomicOr(ws.status, ST_FORMAT);
        if (__any(unsorted) && lane == 0) ws.rowflag[j] = 1;
    }
}

// ---- 3. distinct columns per window ----------------------------------------

struct UniqLds {
    union {
        struct { int32_t key[SORT_CAP]; int32_t idx[SORT_CAP]; } s;
        struct { uint32_t bits[BM_WORDS]; int32_t pre[BM_WORDS]; } b;
    };
    int32_t red[8];
};

__global__ __launch_bounds__(THREADS) void window_unique_kernel(const int32_t *row, const int32_t *col, int64_t nwin,
                                                                int window, Ws ws)
{
    __shared__ UniqLds L;
    const int t = threadIdx.x;
    for (int64_t i = blockIdx.x; i < nwin; i += gridDim.x) {
        if (ws.status[0]) return;
        const int64_t lo = row[i * window], hi = row[i * window + window];
        const int E = int(hi - lo);
        if (E == 0) {
            if (t == 0) ws.nuniq[i] = 0;
            continue;
        }
        int mn = INT_MAX, mx = -1;
        for (int64_t k = lo + t; k < hi; k += THREADS) {
            const int c = col[k];
            mn = min(mn, c);
            mx = max(mx, c);
        }
        block_minmax(mn, mx, L.red);
        const int64_t span = int64_t(mx >> 5) - int64_t(mn >> 5) + 1;   // words
        const bool bitmap = E > SORT_CAP || span <= 8 * int64_t(E);
        if (!bitmap) {
            // Bitonic sort of the window's column ids (padding INT_MAX sorts last).
            int P2 = 64;
            while (P2 < E) P2 <<= 1;
            for (int k = t; k < P2; k += THREADS) {
                L.s.key[k] = k < E ? col[lo + k] : INT_MAX;
                L.s.idx[k] = k;
            }
            __syncthreads();
            for (int size = 2; size <= P2; size <<= 1) {
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    for (int k = t; k < (P2 >> 1); k += THREADS) {
                        const int a = (k / stride) * 2 * stride + (k % stride), b = a + stride;
                        const bool up = (a & size) == 0;
                        const int ka = L.s.key[a], kb = L.s.key[b];
                        if ((ka > kb) == up) {
                            L.s.key[a] = kb; L.s.key[b] = ka;
                            const int ia = L.s.idx[a];
                            L.s.idx[a] = L.s.idx[b]; L.s.idx[b] = ia;
                        }
                    }
                    __syncthreads();
                }
            }
            // Each thread walks a contiguous run of sorted slots.
            const int per = (P2 + THREADS - 1) / THREADS;
            const int k0 = min(t * per, P2), k1 = min(k0 + per, E);
            int cnt = 0;
            for (int k = k0; k < k1; ++k) cnt += (k == 0 || L.s.key[k] != L.s.key[k - 1]);
            int total;
            int r = block_excl_scan(cnt, L.red, total) - 1;
            for (int k = k0; k < k1; ++k) {
                if (k == 0 || L.s.key[k] != L.s.key[k - 1]) {
                    ++r;
                    ws.uniq[lo + r] = L.s.key[k];
                }
                ws.rank[lo + L.s.idx[k]] = r;
            }
            if (t == 0) ws.nuniq[i] = total;
        } else {
            // Bitmap over the window's column span, BM_WORDS words per pass.
            const int64_t w_first = mn >> 5;
            int run = 0;
            for (int64_t c0 = 0; c0 < span; c0 += BM_WORDS) {
                const int64_t w0 = w_first + c0;
                const int nw = int(span - c0 < BM_WORDS ? span - c0 : BM_WORDS);
                for (int w = t; w < nw; w += THREADS) L.b.bits[w] = 0;
                __syncthreads();
                for (int64_t k = lo + t; k < hi; k += THREADS) {
                    const int c = col[k];
                    const int64_t w = int64_t(c >> 5) - w0;
                    if (w >= 0 && w < nw) atomicOr(&L.b.bits[w], 1u << (c & 31));
                }
                __syncthreads();
                constexpr int PER = BM_WORDS / THREADS;
                const int wa = min(t * PER, nw), wb = min(wa + PER, nw);
                int cnt = 0;
                for (int w = wa; w < wb; ++w) cnt += __popc(L.b.bits[w]);
                int total;
                int acc = run + block_excl_scan(cnt, L.red, total);
                for (int w = wa; w < wb; ++w) {
                    L.b.pre[w] = acc;
                    uint32_t bits = L.b.bits[w];
                    while (bits) {
                        const int bit = __ffs(bits) - 1;
                        bits &= bits - 1;
                        ws.uniq[lo + acc] = int32_t((w0 + w) * 32 + bit);
                        ++acc;
                    }
                }
                __syncthreads();
                for (int64_t k = lo + t; k < hi; k += THREADS) {
                    const int c = col[k];
                    const int64_t w = int64_t(c >> 5) - w0;
                    if (w >= 0 && w < nw)
                        ws.rank[k] = L.b.pre[w] + __popc(L.b.bits[w] & ((1u << (c & 31)) - 1u));
                }
                run += total;
                __syncthreads();
            }
            if (t == 0) ws.nuniq[i] = run;
        }
        __syncthreads();
    }
}

// ---- 4. prefix sums over windows (one workgroup) ---------------------------

__global__ __launch_bounds__(THREADS) void window_scan_kernel(int64_t nwin, int window, int wide, int part_size,
                                                              int emit_empty, Ws ws, int64_t *totals)
{
    __shared__ int64_t red[8];
    const int t = threadIdx.x;
    const int st = ws.status[0];
    if (st) {
        if (t == 0) {
            for (int k = 1; k < 8; ++k) totals[k] = 0;
            totals[0] = (st & ST_FORMAT) ? ATTGNN_E_FORMAT : ATTGNN_E_UNSUPPORTED;
        }
        return;
    }
    constexpr int ITEMS = 8;
    int64_t carry_c = 0, carry_p = 0, carry_u = 0;
    for (int64_t base = 0; base < nwin; base += int64_t(THREADS) * ITEMS) {
        int64_t cc[ITEMS], pp[ITEMS];
        int64_t sc = 0, sp = 0, su = 0;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const int64_t i = base + int64_t(t) * ITEMS + k;
            cc[k] = pp[k] = 0;
            if (i < nwin) {
                const int64_t v = ws.nuniq[i];
                const int64_t pad = (v + wide - 1) / wide * wide;
                const int64_t np = part_count_d(pad / wide, part_size, emit_empty);
                cc[k] = np ? pad : 0;
                pp[k] = np;
                su += v;
            }
            sc += cc[k];
            sp += pp[k];
        }
        int64_t tc, tp, tu;
        int64_t oc = block_excl_scan(sc, red, tc) + carry_c;
        int64_t op = block_excl_scan(sp, red, tp) + carry_p;
        block_excl_scan(su, red, tu);
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const int64_t i = base + int64_t(t) * ITEMS + k;
            if (i < nwin) {
                ws.ncol[i] = oc;
                ws.nparts[i] = op;
            }
            oc += cc[k];
            op += pp[k];
        }
        carry_c += tc;
        carry_p += tp;
        carry_u += tu;
    }
    if (t == 0) {
        ws.ncol[nwin] = carry_c;
        ws.nparts[nwin] = carry_p;
        const bool too_big = carry_c > INT32_MAX;
        if (too_big) ws.status[0] = ST_UNSUPPORTED;
        totals[0] = too_big ? ATTGNN_E_UNSUPPORTED : ATTGNN_OK;
        totals[1] = 2 * carry_p + 1;
        totals[2] = carry_c;
        totals[3] = carry_c * window;
        totals[4] = carry_p;
        totals[5] = carry_u;
        totals[6] = 0;
        totals[7] = 0;
    }
}

// ---- 5. fill ---------------------------------------------------------------

__device__ __forceinline__ void store_val(void *base, int64_t idx, int out_type, float v)
{
    switch (out_type) {
    case ATTGNN_F32: static_cast<float *>(base)[idx] = v; break;
    case ATTGNN_F16: static_cast<uint16_t *>(base)[idx] = attgnn::f32_to_f16(v); break;
    default: static_cast<uint16_t *>(base)[idx] = attgnn::f32_to_bf16(v); break;
    }
}

// Elements [s0, s0 + n) of an array of es-byte elements (es = 2 or 4, the
// array es-aligned) set to the 32-bit pattern pat (0 or all ones, the same in
// every byte pair), by the workgroup's THREADS threads: element stores up to
// the first 16-byte boundary and after the last, 16-byte stores between.
__device__ __forceinline__ void fill_range(void *base, int64_t s0, int64_t n, int es, uint32_t pat, int t)
{
    uint8_t *b = static_cast<uint8_t *>(base) + s0 * es;
    const int64_t bytes = n * es;
    int64_t head = int64_t((16 - (reinterpret_cast<uintptr_t>(b) & 15)) & 15);
    if (head > bytes) head = bytes;
    const int64_t nv = (bytes - head) >> 4, tail = head + (nv << 4);
    for (int64_t k = t; k < (head + bytes - tail) / es; k += THREADS) {
        const int64_t off = k < head / es ? k * es : tail + (k - head / es) * es;
        if (es == 4) *reinterpret_cast<uint32_t *>(b + off) = pat;
        else *reinterpret_cast<uint16_t *>(b + off) = uint16_t(pat);
    }
    uint4 *v = reinterpret_cast<uint4 *>(b + head);
    const uint4 p4 = make_uint4(pat, pat, pat, pat);
    for (int64_t k = t; k < nv; k += THREADS) v[k] = p4;
}

// First CSR position of column j in row r (-1 if absent).
__device__ __forceinline__ int32_t find_in_row(const int32_t *row, const int32_t *col, const int32_t *rowflag,
                                               int32_t r, int64_t j)
{
    int64_t b = row[r], e = row[r + 1];
    if (!rowflag[r]) {
        while (b < e) {
            const int64_t m = (b + e) >> 1;
            if (col[m] < j) b = m + 1; else e = m;
        }
        return (b < row[r + 1] && col[b] == j) ? int32_t(b) : -1;
    }
    for (int64_t m = b; m < e; ++m)
        if (col[m] == j) return int32_t(m);
    return -1;
}

struct FillArgs {
    const int32_t *row, *col;
    int64_t nwin;
    int src, out_type, window, wide, part_size;
    const void *src_ptr;
    int32_t *rp, *col_out, *map_t, *win, *atomic;
    void *val;
};

__global__ __launch_bounds__(THREADS) void window_fill_kernel(FillArgs a, Ws ws)
{
    const int t = threadIdx.x;
    if (a.rp && blockIdx.x == 0 && t == 0) a.rp[0] = 0;
    const int W = a.window, wide = a.wide;
    const bool is_map = a.src == ATTGNN_SRC_MAP;
    for (int64_t i = blockIdx.x; i < a.nwin; i += gridDim.x) {
        if (ws.status[0]) return;
        const int64_t p0 = ws.nparts[i], np = ws.nparts[i + 1] - p0;
        if (np == 0) continue;
        const int64_t c0 = ws.ncol[i], pad = ws.ncol[i + 1] - c0;
        const int64_t v = ws.nuniq[i];
        const int64_t lo = a.row[i * W];
        if (a.col_out)
            for (int64_t k = t; k < pad; k += THREADS) a.col_out[c0 + k] = k < v ? ws.uniq[lo + k] : -1;
        const int64_t s0 = c0 * W, ns = pad * W;
        if (is_map) {
            fill_range(a.val, s0, ns, 4, 0xFFFFFFFFu, t);
            if (a.map_t) fill_range(a.map_t, s0, ns, 4, 0xFFFFFFFFu, t);
        } else {
            fill_range(a.val, s0, ns, a.out_type == ATTGNN_F32 ? 4 : 2, 0u, t);
        }
        // parts (example.cpp:112-152); every part but a window's last spans
        // part_size * wide columns
        const int64_t blocks = pad / wide;
        for (int64_t k = t; a.rp && k < np; k += THREADS) {
            int64_t real, padc;
            int at;
            if (blocks == 0) { real = 0; padc = 0; at = 0; }
            else if (blocks <= a.part_size) { real = v; padc = pad - v; at = 0; }
            else {
                const int64_t full = blocks / a.part_size, res = blocks % a.part_size, tailpad = pad - v;
                const int64_t span = int64_t(a.part_size) * wide;
                if (k < full) {
                    const bool last = (res == 0 && k == full - 1);
                    real = last ? span - tailpad : span;
                    padc = last ? tailpad : 0;
                } else {
                    real = res * wide - tailpad;
                    padc = tailpad;
                }
                at = 1;
            }
            const int64_t start = c0 + k * int64_t(a.part_size) * wide;
            const int64_t pi = p0 + k;
            a.rp[2 * pi + 1] = int32_t(start + real);
            a.rp[2 * pi + 2] = int32_t(start + real + padc);
            if (a.win) a.win[pi] = int32_t(i);
            if (a.atomic) a.atomic[pi] = at;
        }
        __syncthreads();   // tile initialisation before the scatter
        for (int p = 0; p < W; ++p) {
            const int64_t j = i * W + p;
            const int64_t e = a.row[j + 1];
            for (int64_t m = a.row[j] + t; m < e; m += THREADS) {
                const int32_t c = ws.rank[m];
                const int32_t cj = a.col[m];
                const int64_t o = s0 + int64_t(c / wide) * W * wide + int64_t(p) * wide + c % wide;
                switch (a.src) {
                case ATTGNN_SRC_DEG_F16: {
                    const uint16_t *d = static_cast<const uint16_t *>(a.src_ptr);
                    store_val(a.val, o, a.out_type, attgnn::f16_to_f32(d[j]) * attgnn::f16_to_f32(d[cj]));
                    break;
                }
                case ATTGNN_SRC_DEG_F32: {
                    const float *d = static_cast<const float *>(a.src_ptr);
                    store_val(a.val, o, a.out_type, d[j] * d[cj]);
                    break;
                }
                case ATTGNN_SRC_EDGE_F32: {
                    // duplicate (row, column) entries: the last one wins, as on the host
                    bool last = true;
                    if (!ws.rowflag[j]) {
                        last = (m + 1 == e) || a.col[m + 1] != cj;
                    } else {
                        for (int64_t q = m + 1; q < e; ++q)
                            if (a.col[q] == cj) { last = false; break; }
                    }
                    if (last) store_val(a.val, o, a.out_type, static_cast<const float *>(a.src_ptr)[m]);
                    break;
                }
                default:
                    atomicMax(static_cast<int32_t *>(a.val) + o, int32_t(m));
                    if (a.map_t) a.map_t[o] = find_in_row(a.row, a.col, ws.rowflag, cj, j);
                }
            }
        }
        __syncthreads();
    }
}

// ---- 6/7. compact tiles ----------------------------------------------------

// Bits 0, 4, 8, ..., 60 of x packed into bits 0..15.
__device__ __forceinline__ uint32_t every_fourth_bit(uint64_t x)
{
    x &= 0x1111111111111111ull;
    x = (x | (x >> 3)) & 0x0303030303030303ull;
    x = (x | (x >> 6)) & 0x000F000F000F000Full;
    x = (x | (x >> 12)) & 0x000000FF000000FFull;
    x = (x | (x >> 24)) & 0xFFFFull;
    return uint32_t(x);
}

// Nonzero pattern of the 8 tile entries at val[i..i+8) (16-byte aligned):
// bit c set when entry c is an entry of the graph (ES-byte tiles; an int32
// edge-id map holds ids >= 0, value tiles nonzero bit patterns).
template <int ES>
__device__ __forceinline__ uint32_t nz8(const void *val, int64_t i, bool is_map)
{
    uint32_t nz = 0;
    if constexpr (ES == 2) {
        const uint4 v = *reinterpret_cast<const uint4 *>(static_cast<const uint16_t *>(val) + i);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            nz |= (uint32_t((w[k] & 0xFFFFu) != 0) << (2 * k)) | (uint32_t((w[k] >> 16) != 0) << (2 * k + 1));
    } else {
        const uint4 *p = reinterpret_cast<const uint4 *>(static_cast<const uint32_t *>(val) + i);
        const uint4 v0 = p[0], v1 = p[1];
        const uint32_t w[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) nz |= uint32_t(is_map ? int32_t(w[k]) >= 0 : w[k] != 0) << k;
    }
    return nz;
}

// One wave per 32-column block of window x 32 tiles (16-byte aligned): lane
// (r, q) = (lane / 4, lane % 4) tests the 8 entries of row r, columns
// 8q..8q+7, with one or two vector loads; ballot j of "column 8q + j is an
// entry" holds column 8q+j's row bits at lanes 4r + q, so its row mask is
// every fourth bit from q, and the block's entry count is the ballots'
// popcount.  counts[nb] = 0 so an exclusive scan of counts[0..nb] ends with
// the total.
template <int ES>
__global__ __launch_bounds__(THREADS) void compact_count_kernel(const void *val, int is_map, int64_t nb, int W,
                                                                uint16_t *masks, int64_t *counts)
{
    const int lane = threadIdx.x & 63, r = lane >> 2, q = lane & 3;
    const int64_t nw = int64_t(gridDim.x) * (THREADS / 64);
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[nb] = 0;
    const int64_t first = int64_t(blockIdx.x) * (THREADS / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int64_t blk = first; blk < nb; blk += nw) {
        const uint32_t nz = r < W ? nz8<ES>(val, (blk * W + r) * 32 + 8 * q, is_map != 0) : 0u;
        uint64_t b[8];
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            b[j] = __ballot((nz >> j) & 1u);
            cnt += __popcll(b[j]);
        }
        if (lane < 32) {
            const int j = lane & 7;
            uint64_t x = b[0];
#pragma unroll
            for (int jj = 1; jj < 8; ++jj) x = j == jj ? b[jj] : x;
            masks[blk * 32 + lane] = uint16_t(every_fourth_bit(x >> (lane >> 3)));
        }
        if (lane == 0) counts[blk] = cnt;
    }
}

__global__ void zero_first_kernel(int64_t *a)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) a[0] = 0;
}

// Workspace of the compact count: the block counts, then rocPRIM's scan storage.
size_t compact_scan_temp_bytes(int64_t nb)
{
    size_t bytes = 0;
    const hipError_t e = rocprim::exclusive_scan(nullptr, bytes, static_cast<const int64_t *>(nullptr),
                                                 static_cast<int64_t *>(nullptr), int64_t(0), size_t(nb + 1),
                                                 rocprim::plus<int64_t>());
    return e == hipSuccess ? bytes : size_t(1) << 40;
}

int64_t compact_ws_bytes(int64_t nb) { return (8 * (nb + 1) + 255) / 256 * 256 + int64_t(compact_scan_temp_bytes(nb)) + 256; }

// One wave per block; lane (g, r) = (lane / 16, lane % 16) packs the entries of
// window row r in the 8-column group g, in column order: the group's 8 row
// masks are one 16-byte load, the row's 8 tile entries one (16-bit tiles) or
// two (32-bit) 16-byte loads, and the entries are picked from registers.
template <int ES>
__global__ __launch_bounds__(THREADS) void compact_pack_kernel(const void *val, int64_t nb, int W,
                                                               const uint16_t *masks, const int64_t *block_base,
                                                               void *values, uint16_t *positions)
{
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const int64_t nw = int64_t(gridDim.x) * (THREADS / 64);
    const int64_t first = int64_t(blockIdx.x) * (THREADS / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // the next block's masks and value offset load while this block packs
    uint4 mv_next = make_uint4(0, 0, 0, 0);
    int64_t base_next = 0;
    if (first < nb) {
        mv_next = *reinterpret_cast<const uint4 *>(masks + first * 32 + 8 * g);
        base_next = block_base[first];
    }
    for (int64_t blk = first; blk < nb; blk += nw) {
        const uint4 mv = mv_next;
        const int64_t base = base_next;
        if (blk + nw < nb) {
            mv_next = *reinterpret_cast<const uint4 *>(masks + (blk + nw) * 32 + 8 * g);
            base_next = block_base[blk + nw];
        }
        const uint32_t mw[4] = {mv.x, mv.y, mv.z, mv.w};
        uint32_t bits = 0;   // bit c set: (r, 8g + c) is an entry
#pragma unroll
        for (int c = 0; c < 8; ++c) bits |= ((mw[c >> 1] >> (16 * (c & 1) + r)) & 1u) << c;
        if (r >= W) bits = 0;
        const int cnt = __popc(bits);
        int64_t o = base + (wave_incl_scan(cnt) - cnt);
        if (!bits) continue;   // after the wave-wide scan; nothing else is shared
        const int64_t i = (blk * W + r) * 32 + 8 * g;
        uint32_t e[8];
        if constexpr (ES == 2) {
            const uint4 v = *reinterpret_cast<const uint4 *>(static_cast<const uint16_t *>(val) + i);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int c = 0; c < 8; ++c) e[c] = (w[c >> 1] >> (16 * (c & 1))) & 0xFFFFu;
        } else {
            const uint4 *p = reinterpret_cast<const uint4 *>(static_cast<const uint32_t *>(val) + i);
            const uint4 v0 = p[0], v1 = p[1];
            e[0] = v0.x; e[1] = v0.y; e[2] = v0.z; e[3] = v0.w;
            e[4] = v1.x; e[5] = v1.y; e[6] = v1.z; e[7] = v1.w;
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if ((bits >> c) & 1u) {
                if constexpr (ES == 4) static_cast<uint32_t *>(values)[o] = e[c];
                else static_cast<uint16_t *>(values)[o] = uint16_t(e[c]);
                if (positions) positions[o] = uint16_t(r * 32 + 8 * g + c);
                ++o;
            }
        }
    }
}

// perm[map[s]] = map_t[s] for every slot with map[s] >= 0.
__global__ __launch_bounds__(THREADS) void map_scatter_kernel(const int32_t *map, const int32_t *map_t, int64_t n,
                                                              int64_t nnz, int64_t *perm)
{
    for (int64_t s = int64_t(blockIdx.x) * THREADS + threadIdx.x; s < n; s += int64_t(gridDim.x) * THREADS) {
        const int32_t e = map[s];
        if (e >= 0 && e < nnz) perm[e] = map_t[s];
    }
}

inline int launch_status()
{
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

}  // namespace

extern "C" int64_t attgnn_blocks_device_workspace_bytes(int64_t rows_plus1, int64_t nnz, int window)
{
    if (rows_plus1 < 1 || nnz < 0 || window <= 0) return -1;
    return ws_bytes(rows_plus1 - 1, nnz, window);
}

extern "C" int attgnn_blocks_plan_device(const int32_t *row, int64_t rows_plus1, const int32_t *col, int64_t nnz,
                                         int window, int wide, int part_size, int emit_empty, void *workspace,
                                         int64_t workspace_bytes, int64_t *totals, attgnn_stream_t stream)
{
    if (!row || !totals || !workspace || rows_plus1 < 1 || window <= 0 || wide <= 0 || part_size <= 0 || nnz < 0 ||
        nnz > INT32_MAX || rows_plus1 - 1 > INT32_MAX)
        return ATTGNN_E_ARG;
    if (nnz > 0 && !col) return ATTGNN_E_ARG;
    const int64_t rows = rows_plus1 - 1, nwin = rows / window;
    if (workspace_bytes < ws_bytes(rows, nnz, window)) return ATTGNN_E_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const Ws ws = ws_carve(workspace, rows, nnz, window);
    hipLaunchKernelGGL(init_kernel, dim3(grid_for(rows, THREADS)), dim3(THREADS), 0, s, ws, rows);
    hipLaunchKernelGGL(rows_check_kernel, dim3(grid_for(rows, THREADS)), dim3(THREADS), 0, s, row, rows, nnz, ws);
    hipLaunchKernelGGL(cols_check_kernel, dim3(grid_for(rows, THREADS / 64)), dim3(THREADS), 0, s, row, col, rows,
                       ws);
    if (nwin > 0)
        hipLaunchKernelGGL(window_unique_kernel, dim3(grid_for(nwin)), dim3(THREADS), 0, s, row, col, nwin, window,
                           ws);
    hipLaunchKernelGGL(window_scan_kernel, dim3(1), dim3(THREADS), 0, s, nwin, window, wide, part_size, emit_empty,
                       ws, totals);
    return launch_status();
}

extern "C" int attgnn_blocks_fill_device(const int32_t *row, int64_t rows_plus1, const int32_t *col, int64_t nnz,
                                         int src, const void *src_ptr, int out_type, int window, int wide,
                                         int part_size, const void *workspace, int64_t workspace_bytes,
                                         int32_t *rp, int32_t *col_out, void *val, int32_t *map_t,
                                         int32_t *window_of_part, int32_t *atomic, attgnn_stream_t stream)
{
    if (!row || !workspace || rows_plus1 < 1 || window <= 0 || wide <= 0 || part_size <= 0 || nnz < 0)
        return ATTGNN_E_ARG;
    if (!rp && (window_of_part || atomic)) return ATTGNN_E_ARG;   // the part table comes with rp
    if (src < ATTGNN_SRC_DEG_F16 || src > ATTGNN_SRC_MAP) return ATTGNN_E_ARG;
    const bool is_map = src == ATTGNN_SRC_MAP;
    if (is_map ? out_type != ATTGNN_I32
               : !(out_type == ATTGNN_F32 || out_type == ATTGNN_F16 || out_type == ATTGNN_BF16))
        return ATTGNN_E_ARG;
    if (!is_map && !src_ptr) return ATTGNN_E_ARG;
    if (map_t && !is_map) return ATTGNN_E_ARG;
    const int64_t rows = rows_plus1 - 1, nwin = rows / window;
    if (workspace_bytes < ws_bytes(rows, nnz, window)) return ATTGNN_E_ARG;
    FillArgs a;
    a.row = row; a.col = col; a.nwin = nwin; a.src = src; a.out_type = out_type; a.window = window;
    a.wide = wide; a.part_size = part_size; a.src_ptr = src_ptr; a.rp = rp; a.col_out = col_out;
    a.map_t = map_t; a.win = window_of_part; a.atomic = atomic; a.val = val;
    hipLaunchKernelGGL(window_fill_kernel, dim3(grid_for(nwin)), dim3(THREADS), 0,
                       reinterpret_cast<hipStream_t>(stream), a,
                       ws_carve(const_cast<void *>(workspace), rows, nnz, window));
    return launch_status();
}

extern "C" int64_t attgnn_blocks_compact_workspace_bytes(int64_t n_cols)
{
    if (n_cols < 0 || n_cols % 32) return -1;
    return compact_ws_bytes(n_cols / 32);
}

extern "C" int attgnn_blocks_compact_count_device(const void *val, int tile_type, int64_t n_cols, int window,
                                                  uint16_t *masks, int64_t *block_base, void *workspace,
                                                  int64_t workspace_bytes, attgnn_stream_t stream)
{
    if (n_cols < 0 || n_cols % 32 || window <= 0 || window > 16 || !block_base) return ATTGNN_E_ARG;
    const int es = type_size(tile_type);
    if (!es) return ATTGNN_E_ARG;
    const int64_t nb = n_cols / 32;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (nb == 0) {
        hipLaunchKernelGGL(zero_first_kernel, dim3(1), dim3(64), 0, s, block_base);
        return launch_status();
    }
    if (!val || !masks || !workspace || workspace_bytes < compact_ws_bytes(nb)) return ATTGNN_E_ARG;
    if ((reinterpret_cast<uintptr_t>(val) | reinterpret_cast<uintptr_t>(masks)) & 15) return ATTGNN_E_ARG;
    // block counts into the workspace, then block_base = their exclusive scan
    // over nb + 1 entries (rocPRIM, device-wide), so block_base[nb] = n_values
    int64_t *counts = static_cast<int64_t *>(workspace);
    uint8_t *temp = static_cast<uint8_t *>(workspace) + (8 * (nb + 1) + 255) / 256 * 256;
    size_t tb = compact_scan_temp_bytes(nb);
    const dim3 grid(grid_for(nb, THREADS / 64));
    if (es == 4)
        hipLaunchKernelGGL(compact_count_kernel<4>, grid, dim3(THREADS), 0, s, val, int(tile_type == ATTGNN_I32), nb,
                           window, masks, counts);
    else
        hipLaunchKernelGGL(compact_count_kernel<2>, grid, dim3(THREADS), 0, s, val, 0, nb, window, masks, counts);
    const hipError_t e = rocprim::exclusive_scan(temp, tb, counts, block_base, int64_t(0), size_t(nb + 1),
                                                 rocprim::plus<int64_t>(), s);
    if (e != hipSuccess) return int(e);
    return launch_status();
}

extern "C" int attgnn_blocks_compact_fill_device(const void *val, int tile_type, int64_t n_cols, int window,
                                                 const uint16_t *masks, const int64_t *block_base, void *values,
                                                 uint16_t *positions, attgnn_stream_t stream)
{
    if (n_cols < 0 || n_cols % 32 || window <= 0 || window > 16 || !block_base) return ATTGNN_E_ARG;
    const int es = type_size(tile_type);
    if (!es) return ATTGNN_E_ARG;
    const int64_t nb = n_cols / 32;
    if (nb == 0) return ATTGNN_OK;
    if (!val || !masks || !values) return ATTGNN_E_ARG;
    if ((reinterpret_cast<uintptr_t>(val) | reinterpret_cast<uintptr_t>(masks)) & 15) return ATTGNN_E_ARG;
    const dim3 grid(grid_for(nb, THREADS / 64));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (es == 4)
        hipLaunchKernelGGL(compact_pack_kernel<4>, grid, dim3(THREADS), 0, s, val, nb, window, masks, block_base,
                           values, positions);
    else
        hipLaunchKernelGGL(compact_pack_kernel<2>, grid, dim3(THREADS), 0, s, val, nb, window, masks, block_base,
                           values, positions);
    return launch_status();
}

extern "C" int attgnn_map_transpose_perm_device(const int32_t *map, const int32_t *map_t, int64_t n_val, int64_t nnz,
                                                int64_t *perm, attgnn_stream_t stream)
{
    if (n_val < 0 || nnz < 0 || (n_val > 0 && (!map || !map_t || !perm))) return ATTGNN_E_ARG;
    if (n_val == 0) return ATTGNN_OK;
    hipLaunchKernelGGL(map_scatter_kernel, dim3(grid_for(n_val, THREADS)), dim3(THREADS), 0,
                       reinterpret_cast<hipStream_t>(stream), map, map_t, n_val, nnz, perm);
    return launch_status();
}

// ---------------------------------------------------------------------------
// Canonical SpMM task table of a window-block structure.  The reference splits
// a window into parts of partSize blocks (example.cpp:120-151), and a split
// window's sum is then grouped by that choice.  Here the SpMM works on tasks
// derived from the windows alone: each window's condensed columns [start,
// padded end) are cut into chunks of chunk_blocks * wide columns from the
// window start (at least one task per window, empty windows included), in
// window order, written in the structure's own format (running (real, pad)
// pairs, window, atomic = the window has more than one chunk).  Every sum
// then depends only on the windows and the chunk size, not on how the caller
// split them.  Parts with a negative window are dropped; task slots past the
// total get window -1 and empty column ranges.
namespace {

__global__ void task_count_kernel(const int32_t *rp, const int32_t *win, int64_t parts, int64_t chunk_cols,
                                  int64_t *counts)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p == parts) counts[parts] = 0;
    if (p >= parts) return;
    const int32_t w = win[p];
    int64_t c = 0;
    if (w >= 0 && (p == 0 || win[p - 1] != w)) {
        int64_t q = p;
        while (q + 1 < parts && win[q + 1] == w) ++q;
        const int64_t span = int64_t(rp[2 * q + 2]) - rp[2 * p];
        c = span <= 0 ? 1 : (span + chunk_cols - 1) / chunk_cols;
    }
    counts[p] = c;
}

__global__ void task_write_kernel(const int32_t *rp, const int32_t *win, int64_t parts, int64_t chunk_cols,
                                  const int64_t *counts, const int64_t *first, int32_t *t_rp, int32_t *t_win,
                                  int32_t *t_atomic)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= parts || counts[p] == 0) return;
    const int32_t w = win[p];
    int64_t q = p;
    while (q + 1 < parts && win[q + 1] == w) ++q;
    const int64_t s0 = rp[2 * p], real_end = rp[2 * q + 1], pad_end = rp[2 * q + 2];
    const int64_t c = counts[p];
    for (int64_t k = 0; k < c; ++k) {
        const int64_t t = first[p] + k;
        const int64_t s = s0 + k * chunk_cols;
        const int64_t e = s + chunk_cols < pad_end ? s + chunk_cols : pad_end;
        const int64_t re = real_end < e ? real_end : e;
        t_rp[2 * t] = int32_t(s);
        t_rp[2 * t + 1] = int32_t(re > s ? re : s);
        t_win[t] = w;
        t_atomic[t] = c > 1 ? 1 : 0;
    }
}

// Slots past the total: window -1, empty ranges at the structure's end; and
// the closing entry of the running pairs.
__global__ void task_tail_kernel(const int32_t *rp, int64_t parts, const int64_t *first, int64_t max_tasks,
                                 int32_t *t_rp, int32_t *t_win, int32_t *t_atomic, int64_t *n_tasks)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int32_t end = rp[2 * parts];
    if (t == 0) {
        t_rp[2 * max_tasks] = end;
        *n_tasks = first[parts];
    }
    if (t >= max_tasks || t < first[parts]) return;
    t_rp[2 * t] = end;
    t_rp[2 * t + 1] = end;
    t_win[t] = -1;
    t_atomic[t] = 0;
}

size_t task_scan_temp_bytes(int64_t parts)
{
    size_t bytes = 0;
    const hipError_t e = rocprim::exclusive_scan(nullptr, bytes, static_cast<const int64_t *>(nullptr),
                                                 static_cast<int64_t *>(nullptr), int64_t(0), size_t(parts + 1),
                                                 rocprim::plus<int64_t>());
    return e == hipSuccess ? bytes : size_t(1) << 40;
}

int64_t task_ws_bytes(int64_t parts)
{
    const int64_t a = (8 * (parts + 1) + 255) / 256 * 256;
    return 2 * a + int64_t(task_scan_temp_bytes(parts)) + 256;
}

}  // namespace

extern "C" int64_t attgnn_blocks_tasks_max(int64_t parts, int64_t n_cols, int wide, int chunk_blocks)
{
    if (parts < 0 || n_cols < 0 || wide <= 0 || chunk_blocks <= 0) return -1;
    const int64_t chunk_cols = int64_t(chunk_blocks) * wide;
    return parts + (n_cols + chunk_cols - 1) / chunk_cols;
}

extern "C" int64_t attgnn_blocks_tasks_workspace_bytes(int64_t parts)
{
    if (parts < 0) return -1;
    return task_ws_bytes(parts);
}

extern "C" int attgnn_blocks_tasks_device(const int32_t *rp, const int32_t *window_of_part, int64_t parts,
                                          int64_t n_cols, int wide, int chunk_blocks, int32_t *task_rp,
                                          int32_t *task_window, int32_t *task_atomic, int64_t max_tasks,
                                          void *workspace, int64_t workspace_bytes, int64_t *n_tasks,
                                          attgnn_stream_t stream)
{
    if (parts < 0 || n_cols < 0 || wide <= 0 || chunk_blocks <= 0 || !rp || !task_rp) return ATTGNN_E_ARG;
    if (max_tasks < attgnn_blocks_tasks_max(parts, n_cols, wide, chunk_blocks)) return ATTGNN_E_ARG;
    if (!workspace || !n_tasks || workspace_bytes < task_ws_bytes(parts)) return ATTGNN_E_ARG;
    if ((parts > 0 && !window_of_part) || (max_tasks > 0 && (!task_window || !task_atomic))) return ATTGNN_E_ARG;
    // one thread per part / task slot (no grid-stride loop): below 2^28 each
    if (parts >= (int64_t(1) << 28) || max_tasks >= (int64_t(1) << 28)) return ATTGNN_E_UNSUPPORTED;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t chunk_cols = int64_t(chunk_blocks) * wide;
    // counts[parts + 1] and their exclusive scan first[parts + 1] in the
    // workspace (first[parts] = the task total), then rocPRIM's storage
    const int64_t a = (8 * (parts + 1) + 255) / 256 * 256;
    int64_t *counts = static_cast<int64_t *>(workspace);
    int64_t *first = reinterpret_cast<int64_t *>(static_cast<uint8_t *>(workspace) + a);
    uint8_t *temp = static_cast<uint8_t *>(workspace) + 2 * a;
    size_t tb = task_scan_temp_bytes(parts);
    hipLaunchKernelGGL(task_count_kernel, dim3(grid_for(parts + 1, THREADS)), dim3(THREADS), 0, s, rp,
                       window_of_part, parts, chunk_cols, counts);
    hipError_t e = rocprim::exclusive_scan(temp, tb, counts, first, int64_t(0), size_t(parts + 1),
                                           rocprim::plus<int64_t>(), s);
    if (e != hipSuccess) return int(e);
    hipLaunchKernelGGL(task_write_kernel, dim3(grid_for(parts, THREADS)), dim3(THREADS), 0, s, rp, window_of_part,
                       parts, chunk_cols, counts, first, task_rp, task_window, task_atomic);
    hipLaunchKernelGGL(task_tail_kernel, dim3(grid_for(std::max<int64_t>(max_tasks, 1), THREADS)), dim3(THREADS), 0,
                       s, rp, parts, first, max_tasks, task_rp, task_window, task_atomic, n_tasks);
    return launch_status();
}
