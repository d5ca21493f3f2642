/*
 * Local-field refresh F = Q . J on the MFMA units (gfx950):
 *   F[y][i] = sum_k q[y][k] J[k][i] = sum_k q[y][k] J[i][k]   (J symmetric)
 * the GEMM (m x N) . (N x Np) that seeds the sweep's LDS fields.
 *
 *  - bf16x2: J = hi + lo (two bf16 planes), spins +-1 exact in bf16,
 *    v_mfma_f32_16x16x32_bf16 twice per k-block, fp32 accumulation.
 *  - fp32:   v_mfma_f32_16x16x4_f32, spins as +-1.0f.
 *
 * Tiling: a block of 4 waves owns 16 output columns i; wave w owns the
 * slice rows [64 w, 64 w + 64) (+ 256 per further pass), i.e. four 16 x 16
 * accumulator tiles.  B fragments come straight from J's rows (symmetry makes
 * the k-run of column i contiguous in row i); A fragments are built from the
 * packed spin bits in registers.  The k dimension is assigned to lane groups
 * identically for A and B, so any consistent permutation of k is valid.
 * The k-blocks past N read zero padding (rows of J are padded to Np, and the
 * spin bits past N are zero), so ragged N needs no special case.
 */
#include "common.h"
#include "kernels.h"

namespace sqaod_rocm {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short i16x8 __attribute__((ext_vector_type(8)));

constexpr int kWave = 64;
constexpr int kRowsPerWave = 64; /* four 16-row tiles */
constexpr int kWavesPerBlock = 4;

/* 8 spins (bits b0..b7 of `bits`) as bf16 +-1 */
__device__ __forceinline__ bf16x8 spinsBF16(uint32_t bits) {
    i16x8 v;
#pragma unroll
    for (int t = 0; t < 8; ++t)
        v[t] = ((bits >> t) & 1u) ? (short)0x3F80 : (short)0xBF80;
    return __builtin_bit_cast(bf16x8, v);
}

/* Operand fetch is double-buffered in super-blocks of kDepth k-blocks: the
 * loads of super-block s+1 are issued before the MFMAs of super-block s, so
 * the global-memory latency overlaps the matrix work. */
constexpr int kDepth = 4;

__device__ __forceinline__ uint32_t spinBits(const uint32_t *Qbits, int y, int m, int Nw, int kk, uint32_t mask) {
    return (y < m) ? ((Qbits[(long long)y * Nw + (kk >> 5)] >> (kk & 31)) & mask) : 0u;
}

__global__ __launch_bounds__(kWavesPerBlock *kWave) void fieldRefreshBF16x2Kernel(const uint16_t *Jhi,
                                                                                  const uint16_t *Jlo,
                                                                                  const uint32_t *Qbits, int N,
                                                                                  int Np, int m, int Nw, float *F) {
    constexpr int KB = 32; /* k per block: one 16x16x32 MFMA per plane */
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int col = blockIdx.x * 16 + (lane & 15); /* B column / C column of this lane */
    const int g = lane >> 4;                       /* lane group: k sub-range and C rows */
    const int nK = (N + KB - 1) / KB;              /* k-blocks; k >= N reads zero padding */
    for (int yBase = w * kRowsPerWave; yBase < m; yBase += kWavesPerBlock * kRowsPerWave) {
        f32x4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        const long long rowOff = (long long)col * Np;
        i16x8 bh[2][kDepth], bl[2][kDepth];
        uint32_t qb[2][kDepth][4];
        auto fetch = [&](int buf, int kb0) {
#pragma unroll
            for (int d = 0; d < kDepth; ++d) {
                const int kk = (kb0 + d) * KB + 8 * g; /* this lane's 8 k values */
                const bool live = (kb0 + d) < nK;
                /* J has N rows: padding columns col >= N contribute zero */
                if (live && col < N) {
                    bh[buf][d] = *reinterpret_cast<const i16x8 *>(Jhi + rowOff + kk);
                    bl[buf][d] = *reinterpret_cast<const i16x8 *>(Jlo + rowOff + kk);
                } else {
                    bh[buf][d] = i16x8{0, 0, 0, 0, 0, 0, 0, 0};
                    bl[buf][d] = bh[buf][d];
                }
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    qb[buf][d][t] = live ? spinBits(Qbits, yBase + 16 * t + (lane & 15), m, Nw, kk, 0xffu) : 0u;
            }
        };
        auto compute = [&](int buf) {
#pragma unroll
            for (int d = 0; d < kDepth; ++d) {
                const bf16x8 bhi = __builtin_bit_cast(bf16x8, bh[buf][d]);
                const bf16x8 blo = __builtin_bit_cast(bf16x8, bl[buf][d]);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const bf16x8 a = spinsBF16(qb[buf][d][t]);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bhi, acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, blo, acc[t], 0, 0, 0);
                }
            }
        };
        /* buffers alternate with constant indices (no dynamic register indexing) */
        fetch(0, 0);
        for (int kb0 = 0; kb0 < nK; kb0 += 2 * kDepth) {
            if (kb0 + kDepth < nK)
                fetch(1, kb0 + kDepth);
            compute(0);
            if (kb0 + kDepth >= nK)
                break;
            if (kb0 + 2 * kDepth < nK)
                fetch(0, kb0 + 2 * kDepth);
            compute(1);
        }
        /* C: lane holds column col, rows 4 g .. 4 g + 3 of each tile */
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int y = yBase + 16 * t + 4 * g + r;
                if (y < m)
                    F[(long long)y * Np + col] = acc[t][r];
            }
    }
}

__global__ __launch_bounds__(kWavesPerBlock *kWave) void fieldRefreshF32Kernel(const float *J, const uint32_t *Qbits,
                                                                               int N, int Np, int m, int Nw,
                                                                               float *F) {
    constexpr int KB = 16; /* k per block: four 16x16x4 MFMAs */
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int col = blockIdx.x * 16 + (lane & 15);
    const int g = lane >> 4;
    const int nK = (N + KB - 1) / KB;
    for (int yBase = w * kRowsPerWave; yBase < m; yBase += kWavesPerBlock * kRowsPerWave) {
        f32x4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        const long long rowOff = (long long)col * Np;
        float4 bv[2][kDepth];
        uint32_t qb[2][kDepth][4];
        auto fetch = [&](int buf, int kb0) {
#pragma unroll
            for (int d = 0; d < kDepth; ++d) {
                /* lane group g holds k = kk + s for MFMA step s (same for A and B) */
                const int kk = (kb0 + d) * KB + 4 * g;
                const bool live = (kb0 + d) < nK;
                bv[buf][d] = (live && col < N) ? *reinterpret_cast<const float4 *>(J + rowOff + kk)
                                               : float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    qb[buf][d][t] = live ? spinBits(Qbits, yBase + 16 * t + (lane & 15), m, Nw, kk, 0xfu) : 0u;
            }
        };
        auto compute = [&](int buf) {
#pragma unroll
            for (int d = 0; d < kDepth; ++d) {
                const float b[4] = {bv[buf][d].x, bv[buf][d].y, bv[buf][d].z, bv[buf][d].w};
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int y = yBase + 16 * t + (lane & 15);
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const float a = (y < m) ? (((qb[buf][d][t] >> s) & 1u) ? 1.f : -1.f) : 0.f;
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[s], acc[t], 0, 0, 0);
                    }
                }
            }
        };
        fetch(0, 0);
        for (int kb0 = 0; kb0 < nK; kb0 += 2 * kDepth) {
            if (kb0 + kDepth < nK)
                fetch(1, kb0 + kDepth);
            compute(0);
            if (kb0 + kDepth >= nK)
                break;
            if (kb0 + 2 * kDepth < nK)
                fetch(0, kb0 + 2 * kDepth);
            compute(1);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int y = yBase + 16 * t + 4 * g + r;
                if (y < m)
                    F[(long long)y * Np + col] = acc[t][r];
            }
    }
}

}  // namespace

void fieldRefreshBF16x2(const uint16_t *d_hi, const uint16_t *d_lo, const uint32_t *d_Qbits, int N, int Np, int m,
                        int Nw, float *d_F, hipStream_t stream) {
    fieldRefreshBF16x2Kernel<<<Np / 16, kWavesPerBlock * kWave, 0, stream>>>(d_hi, d_lo, d_Qbits, N, Np, m, Nw, d_F);
    HIP_CHECK(hipGetLastError());
}

void fieldRefreshF32MFMA(const float *d_J, const uint32_t *d_Qbits, int N, int Np, int m, int Nw, float *d_F,
                         hipStream_t stream) {
    fieldRefreshF32Kernel<<<Np / 16, kWavesPerBlock * kWave, 0, stream>>>(d_J, d_Qbits, N, Np, m, Nw, d_F);
    HIP_CHECK(hipGetLastError());
}

}  // namespace sqaod_rocm
