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

}  // namespace sqaod_rocm
