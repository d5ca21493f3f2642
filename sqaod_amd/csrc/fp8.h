/*
 * OCP FP8 E4M3 conversions of the fp8 precision form (the quantization of J
 * and the decode checks).  Plain C++ with __host__ __device__ qualifiers:
 * include it after <hip/hip_runtime.h> in HIP sources; a host-only build
 * defines the two qualifiers empty.
 */
#pragma once

#include <cstdint>

namespace sqaod_rocm {

/* OCP FP8 E4M3 (e4m3fn: bias 7, 3 mantissa bits, max 448, no infinities),
 * the format of gfx950's FP8 data.  fp8ToFloat is exact; floatToFp8 rounds to
 * nearest even and expects |v| <= 448 (callers scale into range). */
__host__ __device__ inline float fp8ToFloat(uint8_t v) {
    const uint32_t sign = (uint32_t)(v & 0x80u) << 24;
    const uint32_t e = (v >> 3) & 0xFu, mnt = v & 7u;
    float f;
    if (e == 0) {
        f = (float)mnt * 0.001953125f; /* mnt * 2^-9 */
    } else {
        const uint32_t b = ((e + 120u) << 23) | (mnt << 20);
        f = __builtin_bit_cast(float, b);
    }
    return sign ? -f : f;
}

__host__ __device__ inline uint8_t floatToFp8(float v) {
    uint32_t b = __builtin_bit_cast(uint32_t, v);
    const uint8_t sign = (uint8_t)((b >> 24) & 0x80u);
    b &= 0x7fffffffu;
    const float a = __builtin_bit_cast(float, b);
    if (a < 0.015625f) { /* below 2^-6: subnormal quantum 2^-9 (8 rounds up to the smallest normal) */
        const float q = __builtin_rintf(a * 512.f);
        return (uint8_t)(sign | (uint8_t)q);
    }
    const uint32_t r = b + 0x7FFFFu + ((b >> 20) & 1u); /* round the mantissa to 3 bits, ties to even */
    const uint32_t e = (r >> 23) - 120u, mnt = (r >> 20) & 7u;
    if (e > 15u || (e == 15u && mnt == 7u))
        return (uint8_t)(sign | 0x7Eu); /* saturate at 448 */
    return (uint8_t)(sign | (e << 3) | mnt);
}

}  // namespace sqaod_rocm
