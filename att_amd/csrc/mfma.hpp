// gfx950 MFMA operand helpers shared by the SpMM / SDDMM kernels.
//
// Layout of v_mfma_f32_16x16x32_{f16,bf16,fp8_fp8} on a 64-lane wave, with
// g = lane / 16 and n = lane % 16:
//   A operand: lane holds A[m = n][k = 8g .. 8g+7]
//   B operand: lane holds B[k = 8g .. 8g+7][col = n]
//   C/D:       lane holds C[m = 4g + q][col = n], q = 0..3
// v_mfma_f32_16x16x4_f32 uses A[m = n][k = g], B[k = g][col = n] and the same C.
// Only the pairing of A's and B's k slots matters for a contraction, so the
// kernels assign condensed columns to (g, slot) freely.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace attgnn {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ f32x4 mfma_f16(u32x4 a, u32x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_bf16(u32x4 a, u32x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_fp8(u32x2 a, u32x2 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(__builtin_bit_cast(long, a),
                                                      __builtin_bit_cast(long, b), c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_f32(float a, float b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Round-to-nearest-even fp32 -> bf16 bits (finite inputs).
__device__ __forceinline__ uint32_t bf16_bits_rne(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

// Split x = hi + lo with hi, lo bf16 (bf16x3 compensated products).
__device__ __forceinline__ void bf16_split(float x, uint32_t &hi, uint32_t &lo)
{
    hi = bf16_bits_rne(x);
    const float r = x - __uint_as_float(hi << 16);
    lo = bf16_bits_rne(r);
}

// Pack 16-bit element `j` of eight 16-bit vectors into one 8-element operand.
// src[i] holds T 16-bit values packed two per dword.
template <int T>
__device__ __forceinline__ u32x4 gather16(const uint32_t (&src)[8][(T + 1) / 2], int j)
{
    u32x4 r;
    const int w = j >> 1;
    const uint32_t sel = (j & 1) ? 0x07060302u : 0x05040100u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if constexpr (T == 1) {
            r[q] = (src[2 * q][0] & 0xffffu) | (src[2 * q + 1][0] << 16);
        } else {
            r[q] = __builtin_amdgcn_perm(src[2 * q + 1][w], src[2 * q][w], sel);
        }
    }
    return r;
}

// Pack byte `j` of eight byte vectors into one 8-byte operand.
template <int T>
__device__ __forceinline__ u32x2 gather8(const uint32_t (&src)[8][(T + 3) / 4], int j)
{
    u32x2 r;
    const int w = j >> 2, s = (j & 3) * 8;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) v |= ((src[4 * h + b][w] >> s) & 0xffu) << (8 * b);
        r[h] = v;
    }
    return r;
}

}  // namespace attgnn
