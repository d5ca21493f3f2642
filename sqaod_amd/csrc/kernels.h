/*
 * Host-side launchers of the sqaod-rocm HIP kernels (implemented in the .hip
 * sources of this directory).  All launchers are asynchronous on `stream`.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sqaod_rocm {

/* ---------------- random streams (rng.hip) ---------------- */

/* MT19937 state resident in device memory: 624 words and the index of the
 * next word to temper (624 = a twist is due). */
struct MTState {
    uint32_t mt[624];
    uint32_t idx;
};

/* Host-side init_genrand(seed & 0xffffffff) (Random.cpp:70-83). */
void mtSeedHost(MTState *st, unsigned long long seed);

/* Appends n tempered MT19937 outputs to out[0..n), continuing the device
 * state; one workgroup, order identical to the serial host generator. */
void mtGenerate(MTState *d_state, uint32_t *d_out, size_t n, hipStream_t stream);

/* Ordering of attempts inside one sweep, i.e. which stream position feeds
 * attempt (slice y, iteration k). */
enum AttemptOrder {
    orderColoring = 0, /* per iteration k: even slices ascending, then odd (CPUDenseGraphAnnealer.cpp:282-292) */
    orderSliceMajor = 1, /* slice by slice, N attempts each (SA naive, :366-369) */
};

/* Serial-stream mapping: attempt (y, k) consumes words [a*w, a*w + w) with
 * a its position in `order` and w = 2 (float: site, u) or 3 (double: site,
 * 2 draws for u).  Writes sites X[y*N+k] = word % N and uniforms U[y*N+k]. */
template <class real>
void streamToAttempts(const uint32_t *d_words, int N, int m, int order, uint32_t *d_X, real *d_U,
                      hipStream_t stream);

/* Counter-based stream (Philox4x32-10): attempt (y, k) of sweep `sweepIdx`
 * gets one Philox block keyed by the seed. */
template <class real>
void philoxAttempts(unsigned long long seed, unsigned long long sweepIdx, int N, int m, uint32_t *d_X,
                    real *d_U, hipStream_t stream);

/* randomizeSpin from a word stream: bit (y, x) = word[y*N+x] & 1
 * (CPUDenseGraphAnnealer.cpp:152-161: randInt(2) ? +1 : -1).  Qbits is
 * m x Nw words, bit x%32 of word x/32; padding bits are cleared. */
void spinsFromWords(const uint32_t *d_words, int N, int m, int Nw, uint32_t *d_Qbits, hipStream_t stream);
/* randomizeSpin from Philox (domain separated from the attempt streams). */
void spinsFromPhilox(unsigned long long seed, unsigned long long salt, int N, int m, int Nw, uint32_t *d_Qbits,
                     hipStream_t stream);

/* int8 spins (m x N, +-1) <-> packed bits */
void packSpins(const int8_t *d_q, int N, int m, int Nw, uint32_t *d_Qbits, hipStream_t stream);
void unpackSpins(const uint32_t *d_Qbits, int N, int m, int Nw, int8_t *d_q, hipStream_t stream);

/* ---------------- dense local fields (dense_field.hip) ---------------- */

enum FieldFormat {
    fieldF32 = 0,   /* J stored as real */
    fieldBF16x2 = 1 /* J ~ hi + lo, both bf16, hi = bf16(J), lo = bf16(J - hi) */
};

/* Splits J (N x Np, float) into bf16 hi/lo planes of the same shape, and
 * writes Jsplit = float(hi) + float(lo) (exact in fp32): the J the sweep's
 * row updates use, so that fields stay consistent with the split GEMM. */
void splitBF16(const float *d_J, size_t count, uint16_t *d_hi, uint16_t *d_lo, float *d_Jsplit, hipStream_t stream);

/* F[y][i] = sum_j q[y][j] * J[j][i] for all y < m, i < Np (J symmetric, N x Np
 * with zero padding), spins from packed bits.  real = float | double. */
template <class real>
void fieldRefresh(const real *d_J, const uint32_t *d_Qbits, int N, int Np, int m, int Nw, real *d_F,
                  hipStream_t stream);
/* MFMA forms (dense_field_mfma.hip): J given as bf16 hi + lo planes, or fp32
 * J on the fp32 MFMA; fp32 accumulation.  J has N rows, Np columns. */
void fieldRefreshBF16x2Vector(const uint16_t *d_hi, const uint16_t *d_lo, const uint32_t *d_Qbits, int N, int Np,
                              int m, int Nw, float *d_F, hipStream_t stream);
void fieldRefreshF32MFMA(const float *d_J, const uint32_t *d_Qbits, int N, int Np, int m, int Nw, float *d_F,
                         hipStream_t stream);
void fieldRefreshBF16x2(const uint16_t *d_hi, const uint16_t *d_lo, const uint32_t *d_Qbits, int N, int Np, int m,
                        int Nw, float *d_F, hipStream_t stream);

/* ---------------- dense sweep (dense_sweep.hip) ---------------- */

template <class real>
struct SweepArgs {
    int N, Np, m, Nw;
    int sa;                 /* 0: SQA coloring, 1: SA (no Trotter coupling) */
    real twoDivM, coef, beta;
    real invKT;
    const real *J;          /* N x Np rows used by the sweep (J, or its bf16 hi+lo rounding) */
    const real *h;          /* Np */
    real *F;                /* m x Np, in/out */
    uint32_t *Qbits;        /* m x Nw, in/out */
    const uint32_t *X;      /* m x N sites */
    const real *U;          /* m x N uniforms */
    uint32_t *D;            /* m x N decisions (tag << 1 | flipped) */
    uint32_t tag;
    int *abortFlag;         /* set by a wave whose wait exceeds the spin limit */
    unsigned long long *counters; /* [0] accepted flips, [1] batches, [2] blocked waits */
    int spinLimit;
};

/* One annealOneStep sweep: N attempts per slice with exact serial-coloring
 * semantics.  plain = true computes each attempt's field as a J-row . q-row
 * dot (the reference algorithm); plain = false keeps the local fields in LDS.
 * fieldFormat is informational: args.J already holds the rows of the form. */
template <class real>
void denseSweep(const SweepArgs<real> &args, bool plain, int fieldFormat, int slicesPerBlock, hipStream_t stream);

/* Bytes of LDS one block of the local-field sweep needs. */
size_t denseSweepLdsBytes(int Np, int slicesPerBlock, size_t realSize, bool plain);

/* ---------------- energies (energy.hip) ---------------- */

/* FP64 recompute E[y] = -c - h.q_y - q_y^T J q_y (J N x Np real, symmetric
 * with zero diagonal), spins from packed bits; E returned in double. */
template <class real>
void energyFP64(const real *d_J, const real *d_h, double c, const uint32_t *d_Qbits, int N, int Np, int m, int Nw,
                double *d_partial, double *d_E, hipStream_t stream);

/* sum_y q_y . q_(y+1)%m (the Trotter coupling of getSystemE), in double. */
void spinDotSum(const uint32_t *d_Qbits, int N, int m, int Nw, double *d_out, hipStream_t stream);

/* ---------------- self checks (diag.hip) ---------------- */

/* sum over float bit patterns b in [lo, hi) of splitmix64((b << 32) | bits(sq_expf(b))),
 * NaN results canonicalised to 0x7fc00000; computed on the current device. */
unsigned long long expfChecksum(unsigned long long lo, unsigned long long hi);

}  // namespace sqaod_rocm
