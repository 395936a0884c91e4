// Window-block preprocessing (host side of libattgnn.so).
//
// Restates the windowing of gcn/Block_gcn/example.cpp:48-348 (Block_sr_bcrs)
// and the edge-id map of the binary-only blockProcess_sddmm_balance_gnn[_trans]
// (Block/include/example.h:4-6; map layout from
// benchmark-v3/sddmm-benchmark/include/output_tile.h:20,27-33), organised for
// throughput: pass 1 sorts and deduplicates each window's columns in parallel,
// a prefix sum places every window's output, pass 2 fills columns, tiles and
// parts in parallel straight into the final arrays (no per-window maps, no
// critical section, no serial concatenation).  Output is bit-identical to the
// reference for every (window, wide, partSize).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include <omp.h>

#include "attgnn.h"
#include "halfconv.hpp"

struct attgnn_blocks {
    std::vector<int32_t> rp, col, map_t, win, atomic;
    std::vector<uint8_t> val;  // n_val elements of out_type
    int out_type = ATTGNN_F32;
    int window = 0, wide = 0;
    int64_t n_val = 0, n_unique = 0;
    bool has_map_t = false;
};

namespace {

inline size_t type_size(int t)
{
    switch (t) {
    case ATTGNN_F32: case ATTGNN_I32: return 4;
    case ATTGNN_F16: case ATTGNN_BF16: return 2;
    case ATTGNN_FP8E4M3: return 1;
    default: return 0;
    }
}

// Number of (real, pad) pairs a window with `blocks` blocks is split into
// (example.cpp:112-152).
inline int64_t part_count(int64_t blocks, int part_size, bool emit_empty)
{
    if (blocks == 0) return emit_empty ? 1 : 0;
    if (blocks <= part_size) return 1;
    return blocks / part_size + (blocks % part_size ? 1 : 0);
}

inline void store_value(uint8_t *base, int64_t idx, int out_type, float v)
{
    switch (out_type) {
    case ATTGNN_F32: reinterpret_cast<float *>(base)[idx] = v; break;
    case ATTGNN_F16: reinterpret_cast<uint16_t *>(base)[idx] = attgnn::f32_to_f16(v); break;
    case ATTGNN_BF16: reinterpret_cast<uint16_t *>(base)[idx] = attgnn::f32_to_bf16(v); break;
    default: break;
    }
}

}  // namespace

extern "C" const char *attgnn_status_string(int s)
{
    switch (s) {
    case ATTGNN_OK: return "ok";
    case ATTGNN_E_ARG: return "invalid argument";
    case ATTGNN_E_FORMAT: return "malformed sparse structure";
    case ATTGNN_E_NOMEM: return "out of host memory";
    case ATTGNN_E_UNSUPPORTED: return "unsupported configuration";
    case ATTGNN_E_SCRATCH: return "larger than the fixed scratch of the allocation-free entry point";
    default: return "HIP runtime error";
    }
}

extern "C" int attgnn_blocks_build(const int32_t *row, int64_t rows_plus1, const int32_t *col,
                                   int64_t nnz, int src, const void *src_ptr, int out_type,
                                   int window, int wide, int part_size, int emit_empty,
                                   int want_trans, int nthreads, attgnn_blocks **out)
{
    if (!out) return ATTGNN_E_ARG;
    *out = nullptr;
    if (!row || rows_plus1 < 1 || window <= 0 || wide <= 0 || part_size <= 0 || nnz < 0)
        return ATTGNN_E_ARG;
    if (nnz > 0 && !col) return ATTGNN_E_ARG;
    const bool is_map = src == ATTGNN_SRC_MAP;
    if (src < ATTGNN_SRC_DEG_F16 || src > ATTGNN_SRC_MAP) return ATTGNN_E_ARG;
    if (is_map ? out_type != ATTGNN_I32
               : !(out_type == ATTGNN_F32 || out_type == ATTGNN_F16 || out_type == ATTGNN_BF16))
        return ATTGNN_E_ARG;
    if (!is_map && !src_ptr) return ATTGNN_E_ARG;
    if (want_trans && !is_map) return ATTGNN_E_ARG;

    const int64_t rows = rows_plus1 - 1;
    const int64_t nwin = rows / window;  // trailing rows dropped (example.cpp:56-57)
    // CSR sanity: monotone offsets within [0, nnz], column ids in [0, rows).
    if (row[0] != 0 || row[rows] != nnz) return ATTGNN_E_FORMAT;
    for (int64_t i = 0; i < rows; ++i)
        if (row[i + 1] < row[i]) return ATTGNN_E_FORMAT;
    int bad = 0;
#pragma omp parallel for reduction(| : bad) schedule(static)
    for (int64_t m = 0; m < nnz; ++m)
        bad |= (col[m] < 0 || col[m] >= rows);
    if (bad) return ATTGNN_E_FORMAT;

    if (nthreads > 0) omp_set_num_threads(nthreads);
    attgnn_blocks *b = new (std::nothrow) attgnn_blocks();
    if (!b) return ATTGNN_E_NOMEM;
    b->out_type = out_type;
    b->window = window;
    b->wide = wide;

    try {
        // Pass 1: per window, the ascending unique column list, kept in place
        // in a copy of col (a window's union never exceeds its entry count).
        std::vector<int32_t> uniq(col, col + nnz);
        std::vector<int64_t> nuniq(nwin), ncol(nwin + 1, 0), nparts(nwin + 1, 0);
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < nwin; ++i) {
            const int64_t lo = row[i * window], hi = row[i * window + window];
            int32_t *u = uniq.data() + lo;
            std::sort(u, u + (hi - lo));
            const int64_t v = std::unique(u, u + (hi - lo)) - u;
            const int64_t pad = (v + wide - 1) / wide * wide;
            nuniq[i] = v;
            const int64_t np = part_count(pad / wide, part_size, emit_empty != 0);
            ncol[i + 1] = np ? pad : 0;
            nparts[i + 1] = np;
        }
        for (int64_t i = 0; i < nwin; ++i) {
            ncol[i + 1] += ncol[i];
            nparts[i + 1] += nparts[i];
        }
        int64_t U = 0;
        for (int64_t i = 0; i < nwin; ++i) U += nuniq[i];
        const int64_t C = ncol[nwin], P = nparts[nwin];
        if (C > INT32_MAX) {  // rp and col are int32 in the reference format
            delete b;
            return ATTGNN_E_UNSUPPORTED;
        }
        b->n_unique = U;
        b->n_val = C * window;
        b->rp.assign(2 * P + 1, 0);
        b->col.resize(C);
        b->win.resize(P);
        b->atomic.resize(P);
        b->val.assign(static_cast<size_t>(b->n_val) * type_size(out_type), 0);
        if (is_map) {
            int32_t *mv = reinterpret_cast<int32_t *>(b->val.data());
            std::fill(mv, mv + b->n_val, -1);
            if (want_trans) {
                b->map_t.assign(b->n_val, -1);
                b->has_map_t = true;
            }
        }
        // Running offsets: the reference pushes (real, pad) pairs and prefix-sums
        // them (example.cpp:159-170); window i's pairs start at column ncol[i].
        // Pass 2 fills each window independently.
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t i = 0; i < nwin; ++i) {
            const int64_t np = nparts[i + 1] - nparts[i];
            if (np == 0) continue;
            const int64_t v = nuniq[i];
            const int64_t pad = ncol[i + 1] - ncol[i];
            const int64_t c0 = ncol[i];
            const int32_t *u = uniq.data() + row[i * window];
            for (int64_t k = 0; k < pad; ++k) b->col[c0 + k] = k < v ? u[k] : -1;

            const int64_t vbase = c0 * window;
            for (int p = 0; p < window; ++p) {
                const int64_t j = i * window + p;
                for (int64_t m = row[j]; m < row[j + 1]; ++m) {
                    const int64_t c = std::lower_bound(u, u + v, col[m]) - u;
                    const int64_t o = vbase + (c / wide) * window * wide + int64_t(p) * wide + c % wide;
                    switch (src) {
                    case ATTGNN_SRC_DEG_F16: {
                        const uint16_t *d = static_cast<const uint16_t *>(src_ptr);
                        const float prod = attgnn::f16_to_f32(d[j]) * attgnn::f16_to_f32(d[col[m]]);
                        store_value(b->val.data(), o, out_type, prod);
                        break;
                    }
                    case ATTGNN_SRC_DEG_F32: {
                        const float *d = static_cast<const float *>(src_ptr);
                        store_value(b->val.data(), o, out_type, d[j] * d[col[m]]);
                        break;
                    }
                    case ATTGNN_SRC_EDGE_F32:
                        store_value(b->val.data(), o, out_type, static_cast<const float *>(src_ptr)[m]);
                        break;
                    default: {
                        reinterpret_cast<int32_t *>(b->val.data())[o] = static_cast<int32_t>(m);
                        if (want_trans) {
                            // CSR position of (col[m], j): binary search if the row is
                            // sorted, else a scan.
                            const int32_t r = col[m];
                            const int32_t *rb = col + row[r], *re = col + row[r + 1];
                            int32_t pos = -1;
                            if (std::is_sorted(rb, re)) {
                                const int32_t *it = std::lower_bound(rb, re, static_cast<int32_t>(j));
                                if (it != re && *it == j) pos = static_cast<int32_t>(it - col);
                            } else {
                                for (const int32_t *it = rb; it != re; ++it)
                                    if (*it == j) { pos = static_cast<int32_t>(it - col); break; }
                            }
                            b->map_t[o] = pos;
                        }
                    }
                    }
                }
            }

            // Parts (example.cpp:112-152).
            const int64_t blocks = pad / wide;
            // rp[2*pi] (this window's start column, c0) is written by the previous
            // window's last part (or is rp[0] = 0); this window writes rp[2*pi+1..].
            int64_t pi = nparts[i];
            int64_t cur = c0;
            auto emit = [&](int64_t real, int64_t padc, int at) {
                b->rp[2 * pi + 1] = static_cast<int32_t>(cur + real);
                b->rp[2 * pi + 2] = static_cast<int32_t>(cur + real + padc);
                cur += real + padc;
                b->win[pi] = static_cast<int32_t>(i);
                b->atomic[pi] = at;
                ++pi;
            };
            if (blocks <= part_size) {
                emit(v, pad - v, 0);
            } else {
                const int64_t full = blocks / part_size, res = blocks % part_size;
                const int64_t tailpad = pad - v;
                for (int64_t k = 0; k < full; ++k) {
                    const bool last = (res == 0 && k == full - 1);
                    if (last) emit(int64_t(part_size) * wide - tailpad, tailpad, 1);
                    else emit(int64_t(part_size) * wide, 0, 1);
                }
                if (res) emit(res * wide - tailpad, tailpad, 1);
            }
        }
    } catch (const std::bad_alloc &) {
        delete b;
        return ATTGNN_E_NOMEM;
    }
    *out = b;
    return ATTGNN_OK;
}

extern "C" int attgnn_blocks_sizes(const attgnn_blocks *b, int64_t *n_rp, int64_t *n_col,
                                   int64_t *n_val, int64_t *n_parts, int64_t *n_unique,
                                   int *has_map_t)
{
    if (!b) return ATTGNN_E_ARG;
    if (n_rp) *n_rp = static_cast<int64_t>(b->rp.size());
    if (n_col) *n_col = static_cast<int64_t>(b->col.size());
    if (n_val) *n_val = b->n_val;
    if (n_parts) *n_parts = static_cast<int64_t>(b->win.size());
    if (n_unique) *n_unique = b->n_unique;
    if (has_map_t) *has_map_t = b->has_map_t ? 1 : 0;
    return ATTGNN_OK;
}

namespace {

// Copy a result vector out; an empty vector copies nothing (its data() may be
// NULL, which memcpy may not be given even for zero bytes).
template <typename T>
void copy_out(void *dst, const std::vector<T> &v)
{
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
}

}  // namespace

extern "C" int attgnn_blocks_export(const attgnn_blocks *b, int32_t *rp, int32_t *col, void *val,
                                    int32_t *map_t, int32_t *window_of_part, int32_t *atomic)
{
    if (!b) return ATTGNN_E_ARG;
    if (map_t && !b->has_map_t) return ATTGNN_E_ARG;
    copy_out(rp, b->rp);
    copy_out(col, b->col);
    copy_out(val, b->val);
    copy_out(map_t, b->map_t);
    copy_out(window_of_part, b->win);
    copy_out(atomic, b->atomic);
    return ATTGNN_OK;
}

extern "C" void attgnn_blocks_free(attgnn_blocks *b) { delete b; }

// ---------------------------------------------------------------------------
// Compact tiles: per condensed column a 16-bit mask of the window rows holding
// an entry, and the entries' values packed per 32-column block, within a block
// per 8-column group, within a group row-major (row, then column).  A lane that
// owns window row n and one 8-column group then finds its values contiguous at
// block_base[b] + (entries of earlier groups) + (entries of rows < n in its
// group).  An entry is a tile value whose bits are not all zero.
// ---------------------------------------------------------------------------
namespace {

// Entry test: a value tile entry is any nonzero bit pattern; an edge-id map
// entry is an id >= 0 (id 0 is a real edge).
inline bool tile_nonzero(const uint8_t *val, int64_t i, size_t es, bool is_map)
{
    if (is_map) return reinterpret_cast<const int32_t *>(val)[i] >= 0;
    const uint8_t *p = val + i * es;
    for (size_t k = 0; k < es; ++k)
        if (p[k]) return true;
    return false;
}

}  // namespace

extern "C" int attgnn_blocks_compact_sizes(const attgnn_blocks *b, int64_t *n_blocks, int64_t *n_values)
{
    if (!b || b->wide != 32 || b->window > 16) return ATTGNN_E_ARG;
    const int64_t C = static_cast<int64_t>(b->col.size());
    const size_t es = type_size(b->out_type);
    const bool is_map = b->out_type == ATTGNN_I32;
    int64_t nv = 0;
#pragma omp parallel for reduction(+ : nv) schedule(static)
    for (int64_t i = 0; i < b->n_val; ++i) nv += tile_nonzero(b->val.data(), i, es, is_map);
    if (n_blocks) *n_blocks = C / 32;
    if (n_values) *n_values = nv;
    return ATTGNN_OK;
}

extern "C" int attgnn_blocks_compact_export(const attgnn_blocks *b, uint16_t *masks, int64_t *block_base,
                                            void *values, uint16_t *positions)
{
    if (!b || b->wide != 32 || b->window > 16) return ATTGNN_E_ARG;
    if (!masks || !block_base || !values) return ATTGNN_E_ARG;
    const int64_t C = static_cast<int64_t>(b->col.size());
    const int64_t nb = C / 32;
    const int W = b->window;
    const size_t es = type_size(b->out_type);
    const bool is_map = b->out_type == ATTGNN_I32;
    const uint8_t *val = b->val.data();
    std::vector<int64_t> cnt(nb + 1, 0);
#pragma omp parallel for schedule(static)
    for (int64_t blk = 0; blk < nb; ++blk) {
        int64_t n = 0;
        for (int r = 0; r < W; ++r)
            for (int c = 0; c < 32; ++c) {
                const bool nz = tile_nonzero(val, (blk * W + r) * 32 + c, es, is_map);
                n += nz;
            }
        cnt[blk + 1] = n;
        for (int c = 0; c < 32; ++c) {
            uint16_t m = 0;
            for (int r = 0; r < W; ++r)
                if (tile_nonzero(val, (blk * W + r) * 32 + c, es, is_map)) m |= uint16_t(1u << r);
            masks[blk * 32 + c] = m;
        }
    }
    for (int64_t blk = 0; blk < nb; ++blk) cnt[blk + 1] += cnt[blk];
    uint8_t *out = static_cast<uint8_t *>(values);
#pragma omp parallel for schedule(static)
    for (int64_t blk = 0; blk < nb; ++blk) {
        block_base[blk] = cnt[blk];
        int64_t o = cnt[blk];
        for (int g = 0; g < 4; ++g)
            for (int r = 0; r < W; ++r)
                for (int c = 8 * g; c < 8 * g + 8; ++c) {
                    const int64_t i = (blk * W + r) * 32 + c;
                    if (tile_nonzero(val, i, es, is_map)) {
                        std::memcpy(out + o * es, val + i * es, es);
                        if (positions) positions[o] = uint16_t(r * 32 + c);
                        ++o;
                    }
                }
    }
    block_base[nb] = cnt[nb];
    return ATTGNN_OK;
}
