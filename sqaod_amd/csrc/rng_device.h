/*
 * Device-side random helpers shared by the annealers' kernels: the
 * Philox4x32-10 block function (Salmon et al., SC'11) and the word -> uniform
 * conversions of sqaodc/common/Random.cpp:157-168 (float u = float(word) *
 * 2^-32 with round to nearest even, so u may equal 1.0f; double u =
 * ((w1 >> 5) * 2^26 + (w2 >> 6)) * 2^-53).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sqaod_rocm {

/* Philox4x32-10 */
struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
        uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

__device__ __forceinline__ float wordToUniform(uint32_t w) {
    return (float)w * (float)(1. / 4294967296.);
}

__device__ __forceinline__ double wordsToUniform(uint32_t w1, uint32_t w2) {
    uint32_t a = w1 >> 5, b = w2 >> 6;
    return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}


}  // namespace sqaod_rocm
