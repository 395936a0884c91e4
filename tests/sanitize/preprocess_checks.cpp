// Host window-block preprocessing (att_amd/csrc/preprocess.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Built by tests/test_preprocess_sanitized.py with g++ -fsanitize=address,undefined
// together with preprocess.cpp, and run on the CPU.  It drives the C ABI over
// the reference's toy CSR (gcn/SpMM/test/mgcn/test_balance.py:20-22), seeded
// random graphs with hub and empty rows, empty and single-entry graphs, every
// value source and output type, both thread counts, and malformed inputs, and
// checks the structure's invariants:
//   * rp holds running (start, end-of-real) pairs; each part's real columns
//     ascend, its padding is -1 up to a multiple of wide;
//   * every CSR entry of a covered row appears in exactly one tile slot of its
//     window, at its column, with its value (the CSR position for the map
//     source, whose transposed map points at the entry's transpose);
//   * compact tiles: row masks match the nonzero slots, block offsets are their
//     running popcounts, values and positions come in slot order;
//   * malformed CSR -> ATTGNN_E_FORMAT, bad arguments -> ATTGNN_E_ARG, no crash.
// Prints "OK <cases>" and returns 0, or the first failure and returns 1.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "attgnn.h"
#include "halfconv.hpp"

namespace {

int g_cases = 0;

#define CHECK(cond, ...)                                                  \
    do {                                                                  \
        if (!(cond)) {                                                    \
            std::fprintf(stderr, "FAIL %s:%d: %s | ", __FILE__, __LINE__, #cond); \
            std::fprintf(stderr, __VA_ARGS__);                            \
            std::fprintf(stderr, "\n");                                   \
            return false;                                                 \
        }                                                                 \
    } while (0)

struct Csr {
    std::vector<int32_t> row, col;
    int64_t rows() const { return int64_t(row.size()) - 1; }
    int64_t nnz() const { return int64_t(col.size()); }
};

uint64_t lcg(uint64_t &s)
{
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return s >> 33;
}

// Sorted, duplicate-free columns per row; `hubs` rows with n/2 draws.
Csr random_csr(int64_t n, int max_deg, int hubs, uint64_t seed)
{
    Csr c;
    c.row.push_back(0);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t k = i < hubs ? n / 2 : int64_t(lcg(seed) % uint64_t(max_deg + 1));
        std::vector<char> seen(static_cast<size_t>(n), 0);
        for (int64_t t = 0; t < k; ++t) seen[size_t(lcg(seed) % uint64_t(n))] = 1;
        for (int64_t j = 0; j < n; ++j)
            if (seen[size_t(j)]) c.col.push_back(int32_t(j));
        c.row.push_back(int32_t(c.col.size()));
    }
    return c;
}

Csr toy()
{
    Csr c;
    c.row = {0, 6, 7, 8, 10, 10, 11, 13, 15, 18, 19, 19, 19, 20, 20, 20, 20, 24,
             25, 26, 28, 28, 29, 31, 33, 36, 37, 37, 37, 38, 38, 38, 38};
    c.col = {1, 2, 3, 11, 16, 21, 1, 1, 2, 24, 25, 16, 22, 0, 27, 0, 4, 25, 9, 27, 1, 3, 11, 21, 1, 1, 2,
             24, 25, 16, 22, 0, 27, 0, 4, 25, 9, 27};
    return c;
}

// Expected tile value of CSR entry m = (i, j) as a double, per source.
double expected(int src, int out_type, const std::vector<float> &vals, int64_t i, int64_t j, int64_t m)
{
    using namespace attgnn;
    float v;
    if (src == ATTGNN_SRC_DEG_F16) {
        const float a = f16_to_f32(f32_to_f16(vals[size_t(i)]));
        const float b = f16_to_f32(f32_to_f16(vals[size_t(j)]));
        v = a * b;
    } else if (src == ATTGNN_SRC_DEG_F32) {
        v = vals[size_t(i)] * vals[size_t(j)];
    } else {
        v = vals[size_t(m)];
    }
    if (out_type == ATTGNN_F16) return f16_to_f32(f32_to_f16(v));
    if (out_type == ATTGNN_BF16) return bf16_to_f32(f32_to_bf16(v));
    return v;
}

double tile_value(const std::vector<uint8_t> &val, int out_type, int64_t slot)
{
    using namespace attgnn;
    if (out_type == ATTGNN_F32) {
        float f;
        std::memcpy(&f, val.data() + 4 * slot, 4);
        return f;
    }
    uint16_t h;
    std::memcpy(&h, val.data() + 2 * slot, 2);
    return out_type == ATTGNN_F16 ? f16_to_f32(h) : bf16_to_f32(h);
}

int esize(int t) { return t == ATTGNN_F32 || t == ATTGNN_I32 ? 4 : 2; }

bool check_build(const Csr &c, int src, int out_type, int window, int wide, int ps, int emit_empty, int nthreads,
                 const std::vector<float> &vals)
{
    ++g_cases;
    const bool is_map = src == ATTGNN_SRC_MAP;
    attgnn_blocks *b = nullptr;
    // degrees for the fp16 source are passed as fp16 bits
    std::vector<uint16_t> half(vals.size());
    for (size_t k = 0; k < vals.size(); ++k) half[k] = attgnn::f32_to_f16(vals[k]);
    const void *sp = is_map ? nullptr
                            : src == ATTGNN_SRC_DEG_F16 ? static_cast<const void *>(half.data())
                                                        : static_cast<const void *>(vals.data());
    int rc = attgnn_blocks_build(c.row.data(), int64_t(c.row.size()), c.col.empty() ? nullptr : c.col.data(), c.nnz(),
                                 src, sp, out_type, window, wide, ps, emit_empty, is_map ? 1 : 0, nthreads, &b);
    CHECK(rc == ATTGNN_OK && b, "build rc=%d", rc);
    int64_t n_rp, n_col, n_val, n_parts, n_unique;
    int has_t;
    rc = attgnn_blocks_sizes(b, &n_rp, &n_col, &n_val, &n_parts, &n_unique, &has_t);
    CHECK(rc == ATTGNN_OK, "sizes rc=%d", rc);
    CHECK(n_rp == 2 * n_parts + 1 && n_val == n_col * window && has_t == (is_map ? 1 : 0), "size relations");
    std::vector<int32_t> rp(static_cast<size_t>(n_rp)), colo(static_cast<size_t>(n_col)),
        win(static_cast<size_t>(n_parts)), at(static_cast<size_t>(n_parts));
    std::vector<int32_t> mapt(static_cast<size_t>(is_map ? n_val : 0));
    std::vector<uint8_t> val(size_t(n_val * esize(out_type)) + 1);
    rc = attgnn_blocks_export(b, rp.data(), colo.data(), val.data(), is_map ? mapt.data() : nullptr, win.data(),
                              at.data());
    CHECK(rc == ATTGNN_OK, "export rc=%d", rc);

    // rp running pairs, ascending real columns, -1 padding to a multiple of wide
    CHECK(rp[0] == 0 && rp[size_t(n_rp - 1)] == n_col, "rp ends");
    const int64_t nwin = c.rows() / window;
    std::vector<int64_t> seen_slot(static_cast<size_t>(c.nnz()), -1);
    int64_t unique_sum = 0;
    for (int64_t p = 0; p < n_parts; ++p) {
        const int64_t s = rp[size_t(2 * p)], e = rp[size_t(2 * p + 1)], f = rp[size_t(2 * p + 2)];
        CHECK(s <= e && e <= f && (f - s) % wide == 0 && f - e < wide, "part %lld bounds", (long long)p);
        CHECK(win[size_t(p)] >= 0 && win[size_t(p)] < nwin, "window of part %lld", (long long)p);
        CHECK((f - s) / wide <= ps || ps <= 0, "part size %lld", (long long)p);
        for (int64_t k = s; k < e; ++k) {
            CHECK(colo[size_t(k)] >= 0 && colo[size_t(k)] < c.rows(), "column id");
            if (k > s) CHECK(colo[size_t(k)] > colo[size_t(k - 1)], "ascending columns");
        }
        for (int64_t k = e; k < f; ++k) CHECK(colo[size_t(k)] == -1, "padding");
        unique_sum += e - s;
    }
    CHECK(unique_sum == n_unique, "n_unique %lld vs %lld", (long long)unique_sum, (long long)n_unique);
    // parts per window: windows in order, split (atomic) exactly when a window
    // has several parts, an empty window has a part only with emit_empty
    std::vector<int64_t> nparts_of(static_cast<size_t>(nwin > 0 ? nwin : 1), 0);
    for (int64_t p = 0; p < n_parts; ++p) {
        if (p > 0) CHECK(win[size_t(p)] >= win[size_t(p - 1)], "windows in order");
        ++nparts_of[size_t(win[size_t(p)])];
    }
    for (int64_t p = 0; p < n_parts; ++p)
        CHECK(at[size_t(p)] == (nparts_of[size_t(win[size_t(p)])] > 1 ? 1 : 0), "atomic of part %lld",
              (long long)p);
    for (int64_t w = 0; w < nwin; ++w) {
        const bool has = c.row[size_t(w * window + window)] > c.row[size_t(w * window)];
        if (has) CHECK(nparts_of[size_t(w)] >= 1, "window %lld without a part", (long long)w);
        else CHECK(nparts_of[size_t(w)] == (emit_empty ? 1 : 0), "empty window %lld", (long long)w);
    }
    // every slot: either empty or exactly one CSR entry with its value
    for (int64_t p = 0; p < n_parts; ++p) {
        const int64_t w = win[size_t(p)];
        for (int64_t k = rp[size_t(2 * p)]; k < rp[size_t(2 * p + 2)]; ++k) {
            const int64_t blk = k / wide, cb = k % wide;
            for (int r = 0; r < window; ++r) {
                const int64_t slot = (blk * window + r) * wide + cb;
                const int64_t i = w * window + r;
                const int32_t j = colo[size_t(k)];
                // the CSR entry (i, j), if any
                int64_t m = -1;
                if (j >= 0)
                    for (int64_t q = c.row[size_t(i)]; q < c.row[size_t(i + 1)]; ++q)
                        if (c.col[size_t(q)] == j) m = q;
                if (is_map) {
                    int32_t got;
                    std::memcpy(&got, val.data() + 4 * slot, 4);
                    CHECK(got == m, "map slot %lld: %d vs %lld", (long long)slot, got, (long long)m);
                    if (m >= 0) {
                        // transposed map: the CSR position of (j, i), or -1
                        int64_t mt = -1;
                        for (int64_t q = c.row[size_t(j)]; q < c.row[size_t(j + 1)]; ++q)
                            if (c.col[size_t(q)] == i) mt = q;
                        CHECK(mapt[size_t(slot)] == mt, "map_t slot %lld", (long long)slot);
                    }
                } else {
                    const double got = tile_value(val, out_type, slot);
                    const double want = m >= 0 ? expected(src, out_type, vals, i, j, m) : 0.0;
                    CHECK(got == want, "value slot %lld: %g vs %g", (long long)slot, got, want);
                }
                if (m >= 0) {
                    CHECK(seen_slot[size_t(m)] == -1, "entry %lld in two slots", (long long)m);
                    seen_slot[size_t(m)] = slot;
                }
            }
        }
    }
    for (int64_t i = 0; i < nwin * window; ++i)
        for (int64_t q = c.row[size_t(i)]; q < c.row[size_t(i + 1)]; ++q)
            CHECK(seen_slot[size_t(q)] >= 0, "entry %lld of row %lld missing", (long long)q, (long long)i);

    // compact tiles
    if (wide == 32 && window <= 16) {
        int64_t nb, nv;
        rc = attgnn_blocks_compact_sizes(b, &nb, &nv);
        CHECK(rc == ATTGNN_OK && nb == n_col / 32, "compact sizes");
        std::vector<uint16_t> masks(static_cast<size_t>(n_col) + 8), pos(static_cast<size_t>(nv) + 1);
        std::vector<int64_t> base(static_cast<size_t>(nb + 1));
        std::vector<uint8_t> cv(size_t(nv * esize(out_type)) + 4);
        rc = attgnn_blocks_compact_export(b, masks.data(), base.data(), cv.data(), pos.data());
        CHECK(rc == ATTGNN_OK, "compact export");
        int64_t o = 0;
        for (int64_t blk = 0; blk < nb; ++blk) {
            CHECK(base[size_t(blk)] == o, "block base %lld", (long long)blk);
            for (int r = 0; r < window; ++r)
                for (int cb = 0; cb < 32; ++cb) {
                    const int64_t slot = (blk * window + r) * 32 + cb;
                    bool nz;
                    if (is_map) {
                        int32_t v;
                        std::memcpy(&v, val.data() + 4 * slot, 4);
                        nz = v >= 0;
                    } else {
                        nz = tile_value(val, out_type, slot) != 0.0;
                    }
                    CHECK(((masks[size_t(blk * 32 + cb)] >> r) & 1) == (nz ? 1 : 0), "mask bit");
                }
            // values come row by row within 8-column groups: group g, row r, column
            for (int g = 0; g < 4; ++g)
                for (int r = 0; r < window; ++r)
                    for (int cc = 0; cc < 8; ++cc) {
                        const int cb = 8 * g + cc;
                        if ((masks[size_t(blk * 32 + cb)] >> r) & 1) {
                            const int64_t slot = (blk * window + r) * 32 + cb;
                            CHECK(std::memcmp(cv.data() + o * esize(out_type),
                                              val.data() + slot * esize(out_type), size_t(esize(out_type))) == 0,
                                  "packed value %lld", (long long)o);
                            CHECK(pos[size_t(o)] == uint16_t(r * 32 + cb), "position %lld", (long long)o);
                            ++o;
                        }
                    }
        }
        CHECK(base[size_t(nb)] == o && o == nv, "value count");
    }
    attgnn_blocks_free(b);
    return true;
}

bool check_errors()
{
    const Csr c = toy();
    std::vector<float> d(64, 1.0f);
    attgnn_blocks *b = nullptr;
    auto build = [&](const std::vector<int32_t> &row, const std::vector<int32_t> &col, int window, int wide,
                     int ps) {
        ++g_cases;
        const int rc = attgnn_blocks_build(row.data(), int64_t(row.size()), col.empty() ? nullptr : col.data(),
                                           int64_t(col.size()), ATTGNN_SRC_DEG_F32, d.data(), ATTGNN_F32, window,
                                           wide, ps, 1, 0, 2, &b);
        if (b) attgnn_blocks_free(b);
        b = nullptr;
        return rc;
    };
    CHECK(build(c.row, c.col, 0, 8, 2) == ATTGNN_E_ARG, "window 0");
    CHECK(build(c.row, c.col, 8, 0, 2) == ATTGNN_E_ARG, "wide 0");
    CHECK(build(c.row, c.col, 8, 8, 0) == ATTGNN_E_ARG, "part size 0");
    std::vector<int32_t> r = c.row;
    r[0] = 1;
    CHECK(build(r, c.col, 8, 8, 2) == ATTGNN_E_FORMAT, "row[0] != 0");
    r = c.row;
    r[5] = r[6] + 1;
    CHECK(build(r, c.col, 8, 8, 2) == ATTGNN_E_FORMAT, "decreasing row pointers");
    r = c.row;
    r.back() += 1;
    CHECK(build(r, c.col, 8, 8, 2) == ATTGNN_E_FORMAT, "row[n] != nnz");
    std::vector<int32_t> cc = c.col;
    cc[3] = -1;
    CHECK(build(c.row, cc, 8, 8, 2) == ATTGNN_E_FORMAT, "negative column");
    cc[3] = 32;
    CHECK(build(c.row, cc, 8, 8, 2) == ATTGNN_E_FORMAT, "column out of range");
    ++g_cases;
    CHECK(attgnn_blocks_build(c.row.data(), int64_t(c.row.size()), c.col.data(), c.nnz(), ATTGNN_SRC_DEG_F32,
                              nullptr, ATTGNN_F32, 8, 8, 2, 1, 0, 1, &b) == ATTGNN_E_ARG && !b, "no values");
    ++g_cases;
    CHECK(attgnn_blocks_build(c.row.data(), int64_t(c.row.size()), c.col.data(), c.nnz(), ATTGNN_SRC_DEG_F32,
                              d.data(), ATTGNN_F32, 8, 8, 2, 1, 1, 1, &b) == ATTGNN_E_ARG && !b, "trans without map");
    ++g_cases;
    CHECK(attgnn_blocks_build(c.row.data(), int64_t(c.row.size()), c.col.data(), c.nnz(), ATTGNN_SRC_MAP,
                              nullptr, ATTGNN_F16, 8, 8, 2, 1, 0, 1, &b) == ATTGNN_E_ARG && !b, "map to f16");
    ++g_cases;
    CHECK(attgnn_blocks_sizes(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) == ATTGNN_E_ARG,
          "sizes of NULL");
    attgnn_blocks_free(nullptr);
    // compact export of a structure with wide != 32
    ++g_cases;
    CHECK(attgnn_blocks_build(c.row.data(), int64_t(c.row.size()), c.col.data(), c.nnz(), ATTGNN_SRC_DEG_F32,
                              d.data(), ATTGNN_F32, 8, 8, 2, 1, 0, 1, &b) == ATTGNN_OK, "valid build");
    int64_t nb, nv;
    CHECK(attgnn_blocks_compact_sizes(b, &nb, &nv) == ATTGNN_E_ARG, "compact with wide 8");
    attgnn_blocks_free(b);
    return true;
}

}  // namespace

int main()
{
    const int srcs[][2] = {{ATTGNN_SRC_DEG_F16, ATTGNN_F16}, {ATTGNN_SRC_DEG_F32, ATTGNN_F32},
                           {ATTGNN_SRC_EDGE_F32, ATTGNN_F16}, {ATTGNN_SRC_EDGE_F32, ATTGNN_BF16},
                           {ATTGNN_SRC_EDGE_F32, ATTGNN_F32}, {ATTGNN_SRC_MAP, ATTGNN_I32}};
    const int shapes[][3] = {{8, 8, 32}, {8, 8, 1}, {8, 4, 2}, {16, 32, 2}, {16, 32, 1000}, {16, 16, 1},
                             {16, 8, 3}, {8, 32, 2}};
    std::vector<Csr> graphs = {toy(), random_csr(333, 24, 3, 11), random_csr(200, 60, 2, 5), random_csr(50, 3, 0, 7)};
    // empty graph, single entry, no rows
    Csr empty;
    empty.row.assign(33, 0);
    graphs.push_back(empty);
    Csr one;
    one.row.assign(17, 0);
    for (size_t i = 4; i < one.row.size(); ++i) one.row[i] = 1;
    one.col = {9};
    graphs.push_back(one);
    Csr none;
    none.row = {0};
    graphs.push_back(none);

    uint64_t seed = 3;
    for (const Csr &c : graphs) {
        const int64_t n = c.rows();
        std::vector<float> dv(static_cast<size_t>(n > 0 ? n : 1)), ev(static_cast<size_t>(c.nnz() > 0 ? c.nnz() : 1));
        for (auto &x : dv) x = 0.2f + float(lcg(seed) % 1000) / 1000.0f;
        for (auto &x : ev) x = float(int64_t(lcg(seed) % 2001) - 1000) / 337.0f;
        for (const auto &sh : shapes)
            for (const auto &so : srcs)
                for (int emit = 0; emit < 2; ++emit) {
                    const bool deg = so[0] == ATTGNN_SRC_DEG_F16 || so[0] == ATTGNN_SRC_DEG_F32;
                    const int threads = (g_cases % 2) ? 1 : 4;
                    if (!check_build(c, so[0], so[1], sh[0], sh[1], sh[2], emit, threads, deg ? dv : ev)) {
                        std::fprintf(stderr, "case: rows=%lld nnz=%lld src=%d out=%d window=%d wide=%d ps=%d emit=%d\n",
                                     (long long)n, (long long)c.nnz(), so[0], so[1], sh[0], sh[1], sh[2], emit);
                        return 1;
                    }
                }
    }
    if (!check_errors()) return 1;
    std::printf("OK %d\n", g_cases);
    return 0;
}
