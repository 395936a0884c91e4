// IEEE half / bfloat16 conversions (round to nearest even), written with
// integer operations so that host code and device code produce the same bits.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define ATTGNN_HD __host__ __device__
#else
#define ATTGNN_HD
#endif

namespace attgnn {

ATTGNN_HD inline float bits_f32(uint32_t u) { return __builtin_bit_cast(float, u); }
ATTGNN_HD inline uint32_t f32_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

// binary16 -> binary32, exact.
ATTGNN_HD inline float f16_to_f32(uint16_t h)
{
    const uint32_t s = uint32_t(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    if (e == 0x1fu) return bits_f32(s | 0x7f800000u | (m << 13));
    if (e != 0) return bits_f32(s | ((e + 112u) << 23) | (m << 13));
    if (m == 0) return bits_f32(s);
    // subnormal: value = m * 2^-24
    const float v = float(m) * bits_f32(0x33800000u);  // 2^-24
    return s ? -v : v;
}

// binary32 -> binary16, round to nearest even; NaN -> quiet NaN.
ATTGNN_HD inline uint16_t f32_to_f16(float f)
{
    const uint32_t x = f32_bits(f);
    const uint16_t s = uint16_t((x >> 16) & 0x8000u);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return uint16_t(s | 0x7e00u);
    if (ax >= 0x477ff000u) return uint16_t(s | 0x7c00u);  // >= 65520 rounds to inf
    if (ax >= 0x38800000u) {                              // normal half range
        // drop 13 mantissa bits with RNE; carries propagate into the exponent
        const uint32_t r = ax + 0x0fffu + ((ax >> 13) & 1u);
        return uint16_t(s | ((r - 0x38000000u) >> 13));
    }
    // subnormal half (or zero): scale so the half ulp (2^-24) becomes the
    // float ulp at exponent 2^-1 .. then let the FPU round: add 0.5 and read bits
    const float magic = bits_f32(0x3f000000u);  // 0.5 : ulp = 2^-24
    const float t = bits_f32(ax) + magic;
    return uint16_t(s | (f32_bits(t) - 0x3f000000u));
}

// binary32 -> bfloat16, round to nearest even; NaN stays NaN.
ATTGNN_HD inline uint16_t f32_to_bf16(float f)
{
    const uint32_t x = f32_bits(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return uint16_t((x >> 16) | 0x40u);
    return uint16_t((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}

ATTGNN_HD inline float bf16_to_f32(uint16_t b) { return bits_f32(uint32_t(b) << 16); }

}  // namespace attgnn
