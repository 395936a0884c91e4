// Attention-probability pieces shared by the edge softmax (edge.hip) and the
// SpMM that rebuilds the probabilities on the fly (spmm.hip), so that both
// compute them with the same instructions and get the same bits.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace attgnn {

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

// Attention probability of edge e (CSR position) in head h:
// exp(LeakyReLU(logit) - offset), then dropout (kept values scaled).
__device__ __forceinline__ float attn_prob(float logit, float offset, float slope, uint32_t thresh, float keep_scale,
                                           uint64_t seed, int64_t h, int64_t e)
{
    float p = __expf(lrelu(logit, slope) - offset);
    if (thresh) p = keep_edge(seed, h, e, thresh) ? p * keep_scale : 0.f;
    return p;
}

}  // namespace attgnn
