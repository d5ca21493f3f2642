/*
 * bf16 helpers shared by the split of J into hi + lo planes and the MFMA
 * fragment builders: hi = bf16(J) and lo = bf16(J - hi), both rounded to
 * nearest even (finite inputs), so hi + lo keeps the leading 16 significant
 * bits of J and float(hi) + float(lo) is exact in fp32.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sqaod_rocm {

__device__ __forceinline__ uint16_t bf16RoundNE(float v) {
    const uint32_t b = __float_as_uint(v);
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf16ToFloat(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ void bf16Split(float v, uint16_t &hi, uint16_t &lo) {
    hi = bf16RoundNE(v);
    lo = bf16RoundNE(v - bf16ToFloat(hi)); /* v - hi is exact */
}

/* v = hi + mid + lo exactly (three 8-bit significands cover fp32's 24; each
 * difference is exact, and the last remainder fits lo) with all three parts
 * normal bf16 numbers when |v| >= 2^-103, and for v = 0: the fp32 form of the
 * field GEMM runs on these three planes. */
__device__ __forceinline__ void bf16Split3(float v, uint16_t &hi, uint16_t &mid, uint16_t &lo) {
    hi = bf16RoundNE(v);
    const float r1 = v - bf16ToFloat(hi);
    mid = bf16RoundNE(r1);
    lo = bf16RoundNE(r1 - bf16ToFloat(mid));
}

}  // namespace sqaod_rocm
