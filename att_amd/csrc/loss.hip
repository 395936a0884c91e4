// Training-step kernels outside the sparse products: the GCN's loss (the
// negative log-likelihood of a row-wise log-softmax and its gradient) and a
// fused ReLU + dropout.
//
// The reference's GCN takes log_softmax of the last layer
// (end2end_ori/gcn_final/mgat_csr/gat_mgnn.py:39) and nll_loss against the
// labels (gat_mgnn.py:66); in PyTorch that is a chain of a dozen small
// kernels forward and backward over [nodes, classes].  Here 16 lanes take
// one selected row (4 rows per wave, each lane every 16th class, so a row is
// read and written in contiguous pieces), reduce its maximum and its sum of
// exponentials across the 16 lanes, and form the row's loss term
// lse - z[label] and the gradient row (softmax - one-hot) / n, all in FP32,
// storing the gradient in the logits' type; the loss terms are summed per
// workgroup and the workgroup sums in a fixed order by one more workgroup,
// so the loss is the same bits run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "attgnn.h"
#include "halfconv.hpp"

namespace {

constexpr int LOSS_THREADS = 256;

__device__ __forceinline__ float load_logit(const void *p, int type, int64_t i)
{
    if (type == ATTGNN_F32) return static_cast<const float *>(p)[i];
    const uint16_t b = static_cast<const uint16_t *>(p)[i];
    return type == ATTGNN_F16 ? attgnn::f16_to_f32(b) : attgnn::bf16_to_f32(b);
}

__device__ __forceinline__ void store_grad(void *p, int type, int64_t i, float v)
{
    if (type == ATTGNN_F32) static_cast<float *>(p)[i] = v;
    else static_cast<uint16_t *>(p)[i] = type == ATTGNN_F16 ? attgnn::f32_to_f16(v) : attgnn::f32_to_bf16(v);
}

// Sum over a workgroup of LOSS_THREADS threads, fixed order.
__device__ float block_sum(float v, float *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    float s = 0.f;
    if (threadIdx.x == 0)
        for (int w = 0; w < LOSS_THREADS / 64; ++w) s += red[w];
    return s;   // valid in thread 0
}

struct LossArgs {
    const void *logits;
    void *grad;                 // NULL: no gradient
    const float *grad_scale;    // NULL: 1
    const int64_t *labels, *rows;
    float *partial;
    int64_t ld, n_sel, n_rows;
    float scale;
    int type, classes;
};

constexpr int ROW_LANES = 16;                           // lanes per row
constexpr int ROWS_PER_BLOCK = LOSS_THREADS / ROW_LANES;

// max / sum over the 16 lanes of a row (xor partners stay inside the group)
__device__ __forceinline__ float row_max(float v)
{
    for (int o = ROW_LANES / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float row_sum(float v)
{
    for (int o = ROW_LANES / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One row on the 16 lanes of its group: lane k holds classes k, k + 16, ...
// (NV of them in registers, so the row is read once; NV = 0: classes beyond
// 64, re-read from memory for each pass).  Returns the row's loss term in
// lane k == 0 of the group.
template <int NV>
__device__ __forceinline__ float loss_row(const LossArgs &a, int64_t i, int k)
{
    const int64_t r = a.rows ? a.rows[i] : i;
    if (r < 0 || r >= a.n_rows) return k == 0 ? NAN : 0.f;   // outside the logits: NaN loss, nothing stored
    const int64_t y = a.labels[r];
    const int64_t base = r * a.ld;
    float z[NV > 0 ? NV : 1];
    float m = -INFINITY;
    if constexpr (NV > 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int c = k + j * ROW_LANES;
            z[j] = c < a.classes ? load_logit(a.logits, a.type, base + c) : -INFINITY;
            m = fmaxf(m, z[j]);
        }
    } else {
        for (int c = k; c < a.classes; c += ROW_LANES) m = fmaxf(m, load_logit(a.logits, a.type, base + c));
    }
    m = row_max(m);
    float s = 0.f;
    if constexpr (NV > 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) s += expf(z[j] - m);   // exp(-inf) = 0 past the classes
    } else {
        for (int c = k; c < a.classes; c += ROW_LANES) s += expf(load_logit(a.logits, a.type, base + c) - m);
    }
    s = row_sum(s);
    // log-sum-exp = m + log(s), kept apart: (z - m) - log(s) loses no bits
    // to the size of m
    const float ls = logf(s);
    const bool ok = y >= 0 && y < a.classes;
    const float term = k == 0 ? (ok ? (m - load_logit(a.logits, a.type, base + y)) + ls : NAN) : 0.f;
    if (a.grad) {
        const float gs = a.scale * (a.grad_scale ? a.grad_scale[0] : 1.f);
        auto put = [&](int c, float zc) {
            const float p = expf((zc - m) - ls);
            store_grad(a.grad, a.type, base + c, ok ? (p - (c == y ? 1.f : 0.f)) * gs : NAN);
        };
        if constexpr (NV > 0) {
#pragma unroll
            for (int j = 0; j < NV; ++j)
                if (k + j * ROW_LANES < a.classes) put(k + j * ROW_LANES, z[j]);
        } else {
            for (int c = k; c < a.classes; c += ROW_LANES) put(c, load_logit(a.logits, a.type, base + c));
        }
    }
    return term;
}

// Row groups in a grid-stride loop (the 16 lanes of a group share every
// branch, so the lane shuffles of a row stay within active lanes).
template <int NV>
__global__ __launch_bounds__(LOSS_THREADS) void nll_log_softmax_kernel(LossArgs a)
{
    __shared__ float red[LOSS_THREADS / 64];
    const int k = threadIdx.x & (ROW_LANES - 1);
    float term = 0.f;
    for (int64_t i = int64_t(blockIdx.x) * ROWS_PER_BLOCK + threadIdx.x / ROW_LANES; i < a.n_sel;
         i += int64_t(gridDim.x) * ROWS_PER_BLOCK)
        term += loss_row<NV>(a, i, k);
    if (a.partial) {
        const float s = block_sum(term, red);
        if (threadIdx.x == 0) a.partial[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(LOSS_THREADS) void loss_total_kernel(const float *partial, int64_t n, float scale,
                                                                   float *loss)
{
    __shared__ float red[LOSS_THREADS / 64];
    float v = 0.f;
    for (int64_t k = threadIdx.x; k < n; k += LOSS_THREADS) v += partial[k];
    const float s = block_sum(v, red);
    if (threadIdx.x == 0) loss[0] = s * scale;
}

constexpr int64_t LOSS_MAX_BLOCKS = 2048;   // workgroups of the row kernel (its partial sums)

// ---------------------------------------------------------------------------
// ReLU + dropout, forward and backward, 8 elements per thread.  The keep
// decision of element i depends only on (seed, i): keep when the top 24 bits
// of lowbias32(lowbias32(seed_lo ^ seed_hi) + i) are at least p * 2^24, with
// lowbias32 the 32-bit integer hash x ^= x >> 16; x *= 0x7feb352d;
// x ^= x >> 15; x *= 0x846ca68b; x ^= x >> 16 (C. Wellons, "Hash function
// prospector", 2018).  Bit b of mask byte j records whether element 8j + b
// passed (positive and kept): the backward needs nothing else.

__host__ __device__ inline uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

struct DropArgs {
    const void *in;
    void *out;
    uint8_t *mask;          // forward: written
    const uint8_t *mask_in; // backward: read
    int64_t n8;             // groups of 8 elements
    float scale;            // 1 / (1 - p), 0 when p >= 1
    uint32_t thresh, key;   // keep when (hash >> 8) >= thresh; key = lowbias32(seed_lo ^ seed_hi)
    int type;
};

template <int TYPE>
__device__ __forceinline__ void load8(const void *p, int64_t g, float (&v)[8])
{
    if constexpr (TYPE == ATTGNN_F32) {
        const float4 *q = static_cast<const float4 *>(p) + 2 * g;
        const float4 a = q[0], b = q[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
        const uint4 u = static_cast<const uint4 *>(p)[g];
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint16_t lo = uint16_t(w[j] & 0xffffu), hi = uint16_t(w[j] >> 16);
            v[2 * j] = TYPE == ATTGNN_F16 ? attgnn::f16_to_f32(lo) : attgnn::bf16_to_f32(lo);
            v[2 * j + 1] = TYPE == ATTGNN_F16 ? attgnn::f16_to_f32(hi) : attgnn::bf16_to_f32(hi);
        }
    }
}

template <int TYPE>
__device__ __forceinline__ void store8(void *p, int64_t g, const float (&v)[8])
{
    if constexpr (TYPE == ATTGNN_F32) {
        float4 *q = static_cast<float4 *>(p) + 2 * g;
        q[0] = make_float4(v[0], v[1], v[2], v[3]);
        q[1] = make_float4(v[4], v[5], v[6], v[7]);
    } else {
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint16_t lo = TYPE == ATTGNN_F16 ? attgnn::f32_to_f16(v[2 * j]) : attgnn::f32_to_bf16(v[2 * j]);
            const uint16_t hi = TYPE == ATTGNN_F16 ? attgnn::f32_to_f16(v[2 * j + 1]) : attgnn::f32_to_bf16(v[2 * j + 1]);
            w[j] = uint32_t(lo) | (uint32_t(hi) << 16);
        }
        static_cast<uint4 *>(p)[g] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

template <int TYPE>
__global__ __launch_bounds__(256) void relu_dropout_fwd_kernel(DropArgs a)
{
    for (int64_t g = int64_t(blockIdx.x) * 256 + threadIdx.x; g < a.n8; g += int64_t(gridDim.x) * 256) {
        float v[8];
        load8<TYPE>(a.in, g, v);
        uint32_t bits = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint32_t h = lowbias32(a.key + uint32_t(g * 8 + b));
            const bool keep = (h >> 8) >= a.thresh;
            const bool pass = keep && v[b] > 0.f;
            bits |= uint32_t(pass) << b;
            // a kept NaN stays NaN (as relu and dropout leave it); its
            // gradient is 0 (relu's backward passes only positive inputs)
            v[b] = pass || (keep && v[b] != v[b]) ? v[b] * a.scale : 0.f;
        }
        store8<TYPE>(a.out, g, v);
        a.mask[g] = uint8_t(bits);
    }
}

template <int TYPE>
__global__ __launch_bounds__(256) void relu_dropout_bwd_kernel(DropArgs a)
{
    for (int64_t g = int64_t(blockIdx.x) * 256 + threadIdx.x; g < a.n8; g += int64_t(gridDim.x) * 256) {
        float v[8];
        load8<TYPE>(a.in, g, v);
        const uint32_t bits = a.mask_in[g];
#pragma unroll
        for (int b = 0; b < 8; ++b) v[b] = (bits >> b) & 1u ? v[b] * a.scale : 0.f;
        store8<TYPE>(a.out, g, v);
    }
}

int drop_setup(const void *in, int type, int64_t n, float p, const void *out, DropArgs &a)
{
    if (!(type == ATTGNN_F32 || type == ATTGNN_F16 || type == ATTGNN_BF16)) return ATTGNN_E_ARG;
    if (n < 0 || n % 8 != 0 || !(p >= 0.f && p <= 1.f)) return ATTGNN_E_ARG;
    if (n > 0 && (!in || !out)) return ATTGNN_E_ARG;
    if (reinterpret_cast<uintptr_t>(in) % 16 || reinterpret_cast<uintptr_t>(out) % 16) return ATTGNN_E_ARG;
    a.in = in; a.out = const_cast<void *>(out); a.n8 = n / 8; a.type = type;
    a.scale = p >= 1.f ? 0.f : 1.f / (1.f - p);
    // keep when the 24-bit draw is at least p * 2^24 (p = 1: never)
    a.thresh = p >= 1.f ? (1u << 24) : uint32_t(p * 16777216.0f);
    return ATTGNN_OK;
}

unsigned drop_grid(int64_t n8) { return unsigned(std::clamp<int64_t>((n8 + 255) / 256, 1, 8192)); }

}  // namespace

extern "C" int attgnn_relu_dropout_fwd(const void *x, int type, int64_t n, float p, uint64_t seed, void *y,
                                       uint8_t *mask, attgnn_stream_t stream)
{
    DropArgs a{};
    const int rc = drop_setup(x, type, n, p, y, a);
    if (rc != ATTGNN_OK) return rc;
    if (n > 0 && !mask) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    a.mask = mask;
    a.key = lowbias32(uint32_t(seed) ^ uint32_t(seed >> 32));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid{drop_grid(a.n8)}, block{256u};
    if (type == ATTGNN_F16) hipLaunchKernelGGL(relu_dropout_fwd_kernel<ATTGNN_F16>, grid, block, 0, s, a);
    else if (type == ATTGNN_BF16) hipLaunchKernelGGL(relu_dropout_fwd_kernel<ATTGNN_BF16>, grid, block, 0, s, a);
    else hipLaunchKernelGGL(relu_dropout_fwd_kernel<ATTGNN_F32>, grid, block, 0, s, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int attgnn_relu_dropout_bwd(const void *dy, int type, int64_t n, float p, const uint8_t *mask, void *dx,
                                       attgnn_stream_t stream)
{
    DropArgs a{};
    const int rc = drop_setup(dy, type, n, p, dx, a);
    if (rc != ATTGNN_OK) return rc;
    if (n > 0 && !mask) return ATTGNN_E_ARG;
    if (n == 0) return ATTGNN_OK;
    a.mask_in = mask;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid{drop_grid(a.n8)}, block{256u};
    if (type == ATTGNN_F16) hipLaunchKernelGGL(relu_dropout_bwd_kernel<ATTGNN_F16>, grid, block, 0, s, a);
    else if (type == ATTGNN_BF16) hipLaunchKernelGGL(relu_dropout_bwd_kernel<ATTGNN_BF16>, grid, block, 0, s, a);
    else hipLaunchKernelGGL(relu_dropout_bwd_kernel<ATTGNN_F32>, grid, block, 0, s, a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}

extern "C" int64_t attgnn_nll_log_softmax_workspace_elems(int64_t n_sel)
{
    return n_sel < 0 ? -1 : std::clamp<int64_t>((n_sel + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK, 1, LOSS_MAX_BLOCKS);
}

extern "C" int attgnn_nll_log_softmax(const void *logits, int logits_type, int64_t n_rows, int64_t ld, int classes,
                                      const int64_t *labels, const int64_t *rows, int64_t n_sel, float *loss,
                                      void *grad, const float *grad_scale, float *workspace,
                                      int64_t workspace_elems, attgnn_stream_t stream)
{
    if (!(logits_type == ATTGNN_F32 || logits_type == ATTGNN_F16 || logits_type == ATTGNN_BF16)) return ATTGNN_E_ARG;
    if (n_rows < 0 || n_sel < 0 || classes <= 0 || ld < classes || (!loss && !grad)) return ATTGNN_E_ARG;
    if (!rows && n_sel > n_rows) return ATTGNN_E_ARG;
    if (n_sel == 0) return ATTGNN_E_ARG;   // the mean of no rows
    if (!logits || !labels) return ATTGNN_E_ARG;
    if (loss && (!workspace || workspace_elems < attgnn_nll_log_softmax_workspace_elems(n_sel))) return ATTGNN_E_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    LossArgs a;
    a.logits = logits; a.grad = grad; a.grad_scale = grad_scale; a.labels = labels; a.rows = rows;
    a.partial = loss ? workspace : nullptr;
    a.ld = ld; a.n_sel = n_sel; a.n_rows = n_rows; a.scale = 1.0f / float(n_sel); a.type = logits_type; a.classes = classes;
    const int64_t blocks = attgnn_nll_log_softmax_workspace_elems(n_sel);
    const dim3 grid{unsigned(blocks)}, block{unsigned(LOSS_THREADS)};
    if (classes <= 16) hipLaunchKernelGGL(nll_log_softmax_kernel<1>, grid, block, 0, s, a);
    else if (classes <= 32) hipLaunchKernelGGL(nll_log_softmax_kernel<2>, grid, block, 0, s, a);
    else if (classes <= 48) hipLaunchKernelGGL(nll_log_softmax_kernel<3>, grid, block, 0, s, a);
    else if (classes <= 64) hipLaunchKernelGGL(nll_log_softmax_kernel<4>, grid, block, 0, s, a);
    else hipLaunchKernelGGL(nll_log_softmax_kernel<0>, grid, block, 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !loss) return e == hipSuccess ? ATTGNN_OK : int(e);
    hipLaunchKernelGGL(loss_total_kernel, dim3(1), dim3(LOSS_THREADS), 0, s, workspace, blocks, a.scale, loss);
    e = hipGetLastError();
    return e == hipSuccess ? ATTGNN_OK : int(e);
}
