/*
 * sq_expf: a single-precision exp that returns, for every float input, the
 * same bits as the host C library's expf (glibc 2.35, x86-64 FMA variant),
 * so that the device's Metropolis acceptance test `expf(-dE*beta) > u` takes
 * the same branch as the CPU annealer's (CPUDenseGraphAnnealer.cpp:262,
 * `std::exp(-dE * beta)` with real = float).
 *
 * Method: the published table-and-polynomial scheme of ARM's
 * optimized-routines (MIT licence), which glibc's expf also uses:
 *   exp(x) = 2^(k/32) * 2^(r/32),  k = round(x * 32/ln2),  |r| <= 1/2,
 * with 2^(k/32) assembled from a 32-entry table of 2^(i/32) (correctly
 * rounded doubles, minus i<<47 so that the integer part of k can be added to
 * the exponent field directly) and 2^(r/32) from a cubic in r; everything is
 * evaluated in double precision and rounded to float once at the end.
 * The table entries are recomputed from first principles (correctly rounded
 * 2^(i/32) from 80-digit decimal arithmetic) by the CPU test suite.
 *
 * Fused multiply-adds are explicit (__builtin_fma), matching an FMA build of
 * that scheme; every other operation is a single IEEE double operation, so the
 * function is bit-reproducible on any IEEE-754 target (x86-64 host, gfx950).
 * Exhaustive agreement with the host expf over all 2^32 inputs is checked by
 * oracle/check_expf.c.
 */
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define SQ_HD __host__ __device__
#else
#define SQ_HD
#endif

static const uint64_t sq_exp2_tab32[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull,
};

SQ_HD static inline uint32_t sq_f2u(float f) {
    union { float f; uint32_t u; } v; v.f = f; return v.u;
}
SQ_HD static inline uint64_t sq_d2u(double d) {
    union { double d; uint64_t u; } v; v.d = d; return v.u;
}
SQ_HD static inline double sq_u2d(uint64_t u) {
    union { double d; uint64_t u; } v; v.u = u; return v.d;
}

/* Variant switches exist only so that oracle/check_expf.c can show which
 * evaluation order the host library uses; the defaults are the matching one. */
#ifndef SQ_EXPF_FUSED_REDUCTION
#define SQ_EXPF_FUSED_REDUCTION 1
#endif
#ifndef SQ_EXPF_FUSED_POLY
#define SQ_EXPF_FUSED_POLY 1
#endif

/* tab: the 32-entry table above, or a copy of it (e.g. in GPU LDS). */
SQ_HD static inline float sq_expf_tab(float x, const uint64_t *tab) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    /* 32/ln2, 1.5*2^52 and the cubic coefficients of 2^(r/32), scaled by 32^-k */
    const double InvLn2N = 0x1.71547652b82fep+0 * 32.0;
    const double Shift = 0x1.8p+52;
    const double C0 = 0x1.c6af84b912394p-5 / (32.0 * 32.0 * 32.0);
    const double C1 = 0x1.ebfce50fac4f3p-3 / (32.0 * 32.0);
    const double C2 = 0x1.62e42ff0c52d6p-1 / 32.0;

    /* The special inputs are selected after the main path instead of
     * branching before it (a wave whose lanes straddle |x| = 88 would run
     * both sides); the main path is evaluated for every input, with a finite
     * table index, and its result is replaced for x < -103.97 (rounds to +0,
     * also -inf), x > 88.72 (rounds to +inf, also +inf) and NaN (x + x). */
    double xd = (double)x;
#if SQ_EXPF_FUSED_REDUCTION
    double kd = __builtin_fma(InvLn2N, xd, Shift);
    uint64_t ki = sq_d2u(kd);
    kd -= Shift;
    double r = __builtin_fma(InvLn2N, xd, -kd);
#else
    double z = InvLn2N * xd;
    double kd = z + Shift;
    uint64_t ki = sq_d2u(kd);
    kd -= Shift;
    double r = z - kd;
#endif
    uint64_t t = tab[ki % 32];
    t += ki << (52 - 5);
    double s = sq_u2d(t);
    double r2 = r * r;
#if SQ_EXPF_FUSED_POLY
    double p = __builtin_fma(C0, r, C1);
    double y = __builtin_fma(C2, r, 1.0);
    y = __builtin_fma(p, r2, y);
#else
    double p = C0 * r + C1;
    double y = C2 * r + 1.0;
    y = p * r2 + y;
#endif
    y = y * s;
    float res = (float)y;
    res = (x < -0x1.9fe368p6f) ? 0x1p-95f * 0x1p-95f : res;
    res = (x > 0x1.62e42ep6f) ? 0x1p97f * 0x1p97f : res;
    res = (x != x) ? x + x : res;
    return res;
}

SQ_HD static inline float sq_expf(float x) { return sq_expf_tab(x, sq_exp2_tab32); }
