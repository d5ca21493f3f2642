/*
 * Spin-matrix GEMM on the MFMA units (gfx950):
 *   F[y][c] = sum_k q[y][k] B[c][k]
 * with q the +-1 spins of m slices (packed bits, K per row) and B given by
 * rows.  The dense annealer's local-field refresh is B = J (symmetric, so
 * F = Q . J); the bipartite annealer's half-steps are B = J and B = J^T
 * (dEmat = qFixed . J^T, CPUBipartiteGraphAnnealer.cpp:351-352).
 *
 *  - bf16x2: J = hi + lo (two bf16 planes), spins +-1 exact in bf16,
 *    v_mfma_f32_16x16x32_bf16 twice per k-block, fp32 accumulation.
 *  - fp32:   J = hi + mid + lo exactly (three bf16 planes, bf16Split3), the
 *    same kernel with three MFMAs per k-block: every product is exact and
 *    accumulates in fp32, at a third of the v_mfma_f32_16x16x4_f32 cost.
 *  - fp8:    J as e4m3 bytes with a power-of-two scale per column, spins
 *    +-1 exact in e4m3, v_mfma_scale_f32_16x16x128_f8f6f4 with unit block
 *    scales; the column scale multiplies the reduced sum (exactly).
 *
 * J is read from a copy laid out in MFMA fragment order, built once per
 * problem: for column block cb (16 columns) and k-block kb, the 64 lanes'
 * B fragments are one contiguous run (1 KiB per plane), so every operand
 * load of a wave is a single fully coalesced 1 KiB request and a block
 * streams its columns' k-range front to back.
 *
 * Work split: one block of 4 waves owns 16 output columns and all slice rows
 * of a pass (up to 256: sixteen 16 x 16 accumulator tiles per wave).  The
 * k-range is split over the 4 waves in super-blocks of kDepth k-blocks
 * (interleaved), each wave double-buffers its operand loads, and the four
 * partial sums are added in a fixed order through LDS, so results are
 * deterministic.  Each block starts its k-range at a different super-block,
 * so concurrently running blocks read different address ranges.
 *
 * Spins: lane L loads the packed spin words of slice row y0 + 64 r + L for a
 * whole super-block (one 16-byte load per row group); the A fragment of tile
 * t is then fetched from the owning lane with ds_bpermute (__shfl).
 *
 * The k-blocks past the valid length read zero padding (rows of B are padded
 * to K with zeros, the fragment copy is zero for columns past B's rows, and
 * spin bits past the valid length are zero), so ragged shapes need no special
 * case.
 */
#include "common.h"
#include "kernels.h"
#include "bf16.h"

namespace sqaod_rocm {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kWave = 64;
constexpr int kTiles = 16;        /* 16-row accumulator tiles per wave: 256 slice rows per pass */
constexpr int kRowsPerPass = kTiles * 16;
constexpr int kSplitK = 4;        /* waves per block, each a share of the k-range */
constexpr int kDepth = 4;         /* k-blocks per super-block (the double-buffered unit) */
constexpr int kBlockThreads = kSplitK * kWave;

/* 8 spins (bits b0..b7 of `bits`) as bf16 +-1: each dword packs two
 * elements, 0xBF80 (-1) with the sign bit cleared where the spin bit is 1. */
__device__ __forceinline__ bf16x8 spinsBF16(uint32_t bits) {
    uint32_t v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        v[j] = 0xBF80BF80u ^ (((bits >> (2 * j)) & 1u) << 15) ^ ((bits >> (2 * j + 1)) << 31);
    return __builtin_bit_cast(bf16x8, v);
}

/* Fixed-order sum of the kSplitK partial tiles in LDS and the store of the
 * pass's rows: red[w][4 t + r][lane] holds wave w's acc[t][r]; colScale (if
 * given) multiplies each column's sum. */
__device__ __forceinline__ void reduceStore(float (*red)[kTiles * 4][kWave], int y0, int m, int cb, int ldF, float *F,
                                            const float *colScale = nullptr) {
    __syncthreads();
#pragma unroll 4
    for (int j = 0; j < kTiles * 4 * kWave / kBlockThreads; ++j) {
        const int o = j * kBlockThreads + threadIdx.x;
        const int tr = o / kWave, L = o % kWave;
        const int y = y0 + 16 * (tr >> 2) + 4 * (L >> 4) + (tr & 3);
        if (y < m) {
            float s = red[0][tr][L];
#pragma unroll
            for (int w = 1; w < kSplitK; ++w)
                s += red[w][tr][L];
            if (colScale)
                s *= colScale[cb * 16 + (L & 15)];
            F[(long long)y * ldF + cb * 16 + (L & 15)] = s;
        }
    }
    __syncthreads();
}

typedef float RedTile[kTiles * 4][kWave];

/* In a grouped launch the bf16 forms split the waves of a workgroup over
 * columns instead of the k-range when the k-range is at most kColSplitMaxK:
 * many small GEMMs then fill the device with a quarter of the workgroups and
 * no LDS reduction, in the same summation order (refreshBF16Block).  A GEMM
 * alone keeps the k-split, which fills the device better. */
constexpr int kColSplitMaxK = 2048;
__host__ __device__ inline bool bf16ColSplit(int K) { return K <= kColSplitMaxK; }

/* frag: [column blocks][K/32 k-blocks][PLANES][64 lanes] x 16 B (PLANES bf16
 * planes whose sum is B: hi + lo for the bf16x2 form, hi + mid + lo -- exactly
 * B -- for the fp32 form); the block writes columns [16 cb, 16 cb + 16) of F
 * (row stride ldF).  lut: byte -> 8 spins as bf16 +-1, filled by the caller. */
template <int PLANES, bool colSplit = false>
__device__ __forceinline__ void refreshBF16Block(const uint4 *frag, const uint32_t *Qbits, int K, int m, int Nw,
                                                 int ldF, float *F, int cbBlock, RedTile *red, const uint4 *lut) {
    const int lane = threadIdx.x % kWave;
    const int w = threadIdx.x / kWave;
    const int c = lane & 15, g = lane >> 4;
    const int nK = K / 32;      /* k-blocks of 32; K is a multiple of 256 */
    const int nS = nK / kDepth; /* super-blocks */
    /* Without colSplit the four waves take the super-blocks l = w, w + 4, ...
     * of column block cbBlock and their partial tiles are added in LDS in the
     * order ((P0 + P1) + P2) + P3.  With colSplit each wave owns a column block
     * of its own (block cbBlock covers column blocks 4 cbBlock .. 4 cbBlock + 3)
     * and forms the same partials P0..P3 one after the other, adding them in
     * the same order: the results are bit-identical, without the LDS
     * reduction and with a quarter of the workgroups. */
    const int cb = colSplit ? cbBlock * kSplitK + w : cbBlock;
    const int nParts = colSplit ? kSplitK : 1;
    const int sOff = (int)(((long long)cb * 3) % nS);
    const uint4 *fb = frag + (size_t)cb * nK * PLANES * kWave + lane;
    for (int y0 = 0; y0 < m; y0 += kRowsPerPass) {
        f32x4 acc[kTiles], total[kTiles];
      for (int part = 0; part < nParts; ++part) {
        const int lStart = colSplit ? part : w;
        constexpr int lStride = kSplitK;
#pragma unroll
        for (int t = 0; t < kTiles; ++t)
            acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        uint4 bp[PLANES][2][kDepth], qw[2][kTiles / 4];
        /* loads super-block l (logical index) into buffer buf; l < nS */
        auto fetch = [&](int buf, int l) {
            int s = l + sOff;
            if (s >= nS)
                s -= nS;
            const int kb0 = s * kDepth;
#pragma unroll
            for (int d = 0; d < kDepth; ++d) {
                const uint4 *p = fb + (size_t)(kb0 + d) * PLANES * kWave;
#pragma unroll
                for (int pl = 0; pl < PLANES; ++pl)
                    bp[pl][buf][d] = p[pl * kWave];
            }
            /* words kb0 .. kb0+3 of row y0 + 64 r + lane (kb0 is a multiple of 4, Nw of 8);
             * rows past m read row m - 1 and are zeroed, so every load is
             * unconditional and the waits before compute count precisely */
#pragma unroll
            for (int r = 0; r < kTiles / 4; ++r) {
                const int y = y0 + 64 * r + lane;
                const uint4 v = *reinterpret_cast<const uint4 *>(Qbits + (size_t)(y < m ? y : m - 1) * Nw + kb0);
                qw[buf][r] = (y < m) ? v : uint4{0u, 0u, 0u, 0u};
            }
        };
        /* A operands of four tiles are gathered (lane shuffle, then table
         * read) one group ahead of the MFMAs that use them, so a group's LDS
         * round trips run under the previous group's MFMAs; a pass with all
         * 16 tiles in range has no per-tile branch at all */
        const bool fullPass = y0 + kRowsPerPass <= m;
        constexpr int kGroups = kTiles / 4;
        auto groupLive = [&](int t0) { return fullPass || y0 + 16 * t0 < m; }; /* wave-uniform */
        auto gather = [&](int buf, int step, uint4 a4[4]) {
            const int d = step / kGroups, t0 = (step % kGroups) * 4;
            if (!groupLive(t0))
                return;
            const uint4 q4 = qw[buf][t0 >> 2];
            const uint32_t qd = d == 0 ? q4.x : d == 1 ? q4.y : d == 2 ? q4.z : q4.w;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t word = (uint32_t)__shfl((int)qd, 16 * u + c, kWave);
                a4[u] = lut[(word >> (8 * g)) & 0xffu];
            }
        };
        auto compute = [&](int buf) {
            uint4 a4[2][4];
            gather(buf, 0, a4[0]);
#pragma unroll
            for (int step = 0; step < kDepth * kGroups; ++step) {
                if (step + 1 < kDepth * kGroups)
                    gather(buf, step + 1, a4[(step + 1) & 1]);
                const int d = step / kGroups, t0 = (step % kGroups) * 4;
                if (!groupLive(t0))
                    continue;

#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int t = t0 + u;
                    if (fullPass || y0 + 16 * t < m) {
                        const bf16x8 av = __builtin_bit_cast(bf16x8, a4[step & 1][u]);
#pragma unroll
                        for (int pl = 0; pl < PLANES; ++pl)
                            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                av, __builtin_bit_cast(bf16x8, bp[pl][buf][d]), acc[t], 0, 0, 0);
                    }
                }
            }
        };
        /* buffers alternate with constant indices (no dynamic register indexing) */
        for (int l = lStart; l < nS; l += 2 * lStride) {
            if (l == lStart)
                fetch(0, l);
            if (l + lStride < nS)
                fetch(1, l + lStride);
            compute(0);
            if (l + lStride >= nS)
                break;
            if (l + 2 * lStride < nS)
                fetch(0, l + 2 * lStride);
            compute(1);
        }
        if (colSplit) {
#pragma unroll
            for (int t = 0; t < kTiles; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    total[t][r] = part == 0 ? acc[t][r] : total[t][r] + acc[t][r];
        }
      }
        if (colSplit) {
            /* tile t, register r: row y0 + 16 t + 4 (lane / 16) + r, column 16 cb + lane % 16 */
#pragma unroll
            for (int t = 0; t < kTiles; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int y = y0 + 16 * t + 4 * g + r;
                    if (y < m)
                        F[(long long)y * ldF + cb * 16 + c] = total[t][r];
                }
            continue;
        }
#pragma unroll
        for (int t = 0; t < kTiles; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                red[w][4 * t + r][lane] = acc[t][r];
        reduceStore(red, y0, m, cb, ldF, F);
    }
}

__device__ __forceinline__ void fillLutBF16(uint4 *lut) {
    for (int i = threadIdx.x; i < 256; i += kBlockThreads)
        lut[i] = __builtin_bit_cast(uint4, spinsBF16((uint32_t)i));
    __syncthreads();
}

__global__ __launch_bounds__(kBlockThreads) void fieldRefreshBF16x2Kernel(const uint4 *frag, const uint32_t *Qbits,
                                                                          int K, int m, int Nw, int ldF, float *F) {
    __shared__ float red[kSplitK][kTiles * 4][kWave]; /* 64 KiB */
    __shared__ uint4 lut[256];                        /* 4 KiB */
    fillLutBF16(lut);
    refreshBF16Block<2>(frag, Qbits, K, m, Nw, ldF, F, blockIdx.x, red, lut);
}

__global__ __launch_bounds__(kBlockThreads) void fieldRefreshBF16x3Kernel(const uint4 *frag, const uint32_t *Qbits,
                                                                          int K, int m, int Nw, int ldF, float *F) {
    __shared__ float red[kSplitK][kTiles * 4][kWave]; /* 64 KiB */
    __shared__ uint4 lut[256];                        /* 4 KiB */
    fillLutBF16(lut);
    refreshBF16Block<3>(frag, Qbits, K, m, Nw, ldF, F, blockIdx.x, red, lut);
}


typedef int i32x8 __attribute__((ext_vector_type(8)));
constexpr int kDepth8 = 2; /* fp8: k-blocks of 128 per super-block (256 k = 8 spin words) */
constexpr int kE4M3Unit = 127; /* E8M0 block scale 2^0 */

/* 8 spins (bits of `bits`) as e4m3 +-1: 0x38 = +1, 0xB8 = -1 */
__device__ __forceinline__ uint2 spinsFP8(uint32_t bits) {
    uint32_t v[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const uint32_t nib = (bits >> (4 * j)) & 15u;
        v[j] = 0xB8B8B8B8u ^ (((nib * 0x204081u) & 0x01010101u) << 7);
    }
    return uint2{v[0], v[1]};
}

/* frag: [column blocks][K/128 k-blocks][2 halves][64 lanes] x 16 B.  In
 * super-block s (k-blocks 2s, 2s + 1), lane group g of k-block 2s + d holds
 * k = 256 s + 64 g + 32 d + e, e < 32: spin word 8 s + 2 g + d of the row,
 * bit e -- so each lane's A operand for a whole super-block is the 8-byte
 * pair of words (8 s + 2 g, 8 s + 2 g + 1) of its row, loaded directly. */
__device__ __forceinline__ void refreshFP8Block(const uint4 *frag, const float *colScale, const uint32_t *Qbits,
                                                int K, int m, int Nw, int ldF, float *F, int cb, RedTile *red,
                                                const uint2 *lut) {
    const int lane = threadIdx.x % kWave;
    const int w = threadIdx.x / kWave;
    const int c = lane & 15, g = lane >> 4;
    const int nK = K / 128;
    const int nS = nK / kDepth8;
    const int sOff = (int)(((long long)cb * 3) % nS);
    const uint4 *fb = frag + (size_t)cb * nK * 2 * kWave + lane;
    for (int y0 = 0; y0 < m; y0 += kRowsPerPass) {
        f32x4 acc[kTiles];
#pragma unroll
        for (int t = 0; t < kTiles; ++t)
            acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        uint4 b0[2][kDepth8], b1[2][kDepth8];
        uint2 qv[2][kTiles];
        auto fetch = [&](int buf, int l) {
            int s = l + sOff;
            if (s >= nS)
                s -= nS;
#pragma unroll
            for (int d = 0; d < kDepth8; ++d) {
                const uint4 *p = fb + (size_t)(s * kDepth8 + d) * 2 * kWave;
                b0[buf][d] = p[0];
                b1[buf][d] = p[kWave];
            }
#pragma unroll
            for (int t = 0; t < kTiles; ++t) {
                const int y = y0 + 16 * t + c;
                const uint2 v = *reinterpret_cast<const uint2 *>(Qbits + (size_t)(y < m ? y : m - 1) * Nw + 8 * s + 2 * g);
                qv[buf][t] = (y < m) ? v : uint2{0u, 0u};
            }
        };
        /* A operands (four table reads per tile) are gathered one group of
         * four tiles ahead of the MFMAs that use them */
        constexpr int kGroups = kTiles / 4;
        auto gather = [&](int buf, int step, i32x8 aop[4]) {
            const int d = step / kGroups, t0 = (step % kGroups) * 4;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = t0 + u;
                if (y0 + 16 * t < m) { /* wave-uniform */
                    const uint32_t word = d == 0 ? qv[buf][t].x : qv[buf][t].y;
                    const uint2 l0 = lut[word & 0xffu], l1 = lut[(word >> 8) & 0xffu];
                    const uint2 l2 = lut[(word >> 16) & 0xffu], l3 = lut[word >> 24];
                    aop[u] = i32x8{(int)l0.x, (int)l0.y, (int)l1.x, (int)l1.y,
                                   (int)l2.x, (int)l2.y, (int)l3.x, (int)l3.y};
                }
            }
        };
        auto compute = [&](int buf) {
            i32x8 aop[2][4];
            gather(buf, 0, aop[0]);
#pragma unroll
            for (int step = 0; step < kDepth8 * kGroups; ++step) {
                if (step + 1 < kDepth8 * kGroups)
                    gather(buf, step + 1, aop[(step + 1) & 1]);
                const int d = step / kGroups, t0 = (step % kGroups) * 4;
                const uint4 x0 = b0[buf][d], x1 = b1[buf][d];
                const i32x8 bop = {(int)x0.x, (int)x0.y, (int)x0.z, (int)x0.w,
                                   (int)x1.x, (int)x1.y, (int)x1.z, (int)x1.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int t = t0 + u;
                    if (y0 + 16 * t < m) /* wave-uniform */
                        acc[t] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(aop[step & 1][u], bop, acc[t], 0, 0,
                                                                                  0, kE4M3Unit, 0, kE4M3Unit);
                }
            }
        };
        for (int l = w; l < nS; l += 2 * kSplitK) {
            if (l == w)
                fetch(0, l);
            if (l + kSplitK < nS)
                fetch(1, l + kSplitK);
            compute(0);
            if (l + kSplitK >= nS)
                break;
            if (l + 2 * kSplitK < nS)
                fetch(0, l + 2 * kSplitK);
            compute(1);
        }
#pragma unroll
        for (int t = 0; t < kTiles; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                red[w][4 * t + r][lane] = acc[t][r];
        reduceStore(red, y0, m, cb, ldF, F, colScale);
    }
}

__device__ __forceinline__ void fillLutFP8(uint2 *lut) {
    for (int i = threadIdx.x; i < 256; i += kBlockThreads)
        lut[i] = spinsFP8((uint32_t)i);
    __syncthreads();
}

__global__ __launch_bounds__(kBlockThreads) void fieldRefreshFP8Kernel(const uint4 *frag, const float *colScale,
                                                                       const uint32_t *Qbits, int K, int m, int Nw,
                                                                       int ldF, float *F) {
    __shared__ float red[kSplitK][kTiles * 4][kWave]; /* 64 KiB */
    __shared__ uint2 lut[256];                        /* 2 KiB */
    fillLutFP8(lut);
    refreshFP8Block(frag, colScale, Qbits, K, m, Nw, ldF, F, blockIdx.x, red, lut);
}

/* Many GEMMs of one form in one launch: block b runs column block
 * b - blockStart[j] of job j, blockStart[j] <= b < blockStart[j + 1]. */
__device__ __forceinline__ int jobOfBlock(const int *blockStart, int nJobs) {
    int lo = 0, hi = nJobs - 1;
    const int b = blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blockStart[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int FORM>
__global__ __launch_bounds__(kBlockThreads) void fieldRefreshGroupedKernel(const FieldGemmJob *jobs,
                                                                           const int *blockStart, int nJobs) {
    __shared__ float red[kSplitK][kTiles * 4][kWave];
    __shared__ uint4 lut[256];
    const int j = jobOfBlock(blockStart, nJobs);
    const FieldGemmJob jb = jobs[j];
    const int cb = (int)blockIdx.x - blockStart[j];
    if constexpr (FORM == fieldBF16x2) {
        fillLutBF16(lut);
        /* the split as a template argument: each path allocates its own registers */
        if (bf16ColSplit(jb.K))
            refreshBF16Block<2, true>(reinterpret_cast<const uint4 *>(jb.frag), jb.Qbits, jb.K, jb.m, jb.Nw, jb.ldF,
                                      jb.F, cb, red, lut);
        else
            refreshBF16Block<2, false>(reinterpret_cast<const uint4 *>(jb.frag), jb.Qbits, jb.K, jb.m, jb.Nw, jb.ldF,
                                       jb.F, cb, red, lut);
    } else if constexpr (FORM == fieldFP8) {
        fillLutFP8(reinterpret_cast<uint2 *>(lut));
        refreshFP8Block(reinterpret_cast<const uint4 *>(jb.frag), jb.colScale, jb.Qbits, jb.K, jb.m, jb.Nw, jb.ldF,
                        jb.F, cb, red, reinterpret_cast<const uint2 *>(lut));
    } else {
        fillLutBF16(lut);
        if (bf16ColSplit(jb.K))
            refreshBF16Block<3, true>(reinterpret_cast<const uint4 *>(jb.frag), jb.Qbits, jb.K, jb.m, jb.Nw, jb.ldF,
                                      jb.F, cb, red, lut);
        else
            refreshBF16Block<3, false>(reinterpret_cast<const uint4 *>(jb.frag), jb.Qbits, jb.K, jb.m, jb.Nw, jb.ldF,
                                       jb.F, cb, red, lut);
    }
}

/* One thread per (column block, k-block, lane): the 32 bytes of the lane's
 * column (a row of B8, stride ld8, zero past `rows`) in the order above. */
__global__ void buildFragFP8Kernel(const uint8_t *B8, int rows, int ld8, int nColBlocks, int K, uint4 *frag) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nK = K / 128;
    if (idx >= (size_t)nColBlocks * nK * kWave)
        return;
    const int lane = (int)(idx % kWave);
    const size_t rest = idx / kWave;
    const int kb = (int)(rest % nK), cb = (int)(rest / nK);
    const int col = cb * 16 + (lane & 15);
    const int k0 = 256 * (kb / kDepth8) + 64 * (lane >> 4) + 32 * (kb % kDepth8);
    uint4 v0 = {0u, 0u, 0u, 0u}, v1 = {0u, 0u, 0u, 0u};
    if (col < rows) {
        const uint4 *src = reinterpret_cast<const uint4 *>(B8 + (size_t)col * ld8 + k0);
        v0 = src[0];
        v1 = src[1];
    }
    uint4 *dst = frag + ((size_t)cb * nK + kb) * 2 * kWave + lane;
    dst[0] = v0;
    dst[kWave] = v1;
}

/* One thread per (column block, k-block, lane): the lane's 8 k values of its
 * column (a row of B, row stride ld, zero past `rows`), split into bf16 hi and lo. */
template <int PLANES>
__global__ void buildFragBF16Kernel(const float *B, int rows, int ld, int nColBlocks, int K, uint4 *frag) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nK = K / 32;
    if (idx >= (size_t)nColBlocks * nK * kWave)
        return;
    const int lane = (int)(idx % kWave);
    const size_t rest = idx / kWave;
    const int kb = (int)(rest % nK), cb = (int)(rest / nK);
    const int col = cb * 16 + (lane & 15), k0 = kb * 32 + 8 * (lane >> 4);
    uint32_t pw[PLANES][4] = {};
    if (col < rows) {
        const float4 *src = reinterpret_cast<const float4 *>(B + (size_t)col * ld + k0);
        const float4 v0 = src[0], v1 = src[1];
        const float v[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            uint16_t h[3];
            if constexpr (PLANES == 2)
                bf16Split(v[e], h[0], h[1]);
            else
                bf16Split3(v[e], h[0], h[1], h[2]);
#pragma unroll
            for (int pl = 0; pl < PLANES; ++pl)
                pw[pl][e >> 1] |= (uint32_t)h[pl] << (16 * (e & 1));
        }
    }
    uint4 *dst = frag + ((size_t)cb * nK + kb) * PLANES * kWave + lane;
#pragma unroll
    for (int pl = 0; pl < PLANES; ++pl)
        dst[pl * kWave] = uint4{pw[pl][0], pw[pl][1], pw[pl][2], pw[pl][3]};
}

void checkGemm(int nCols, int K, int ld, int Nw) {
    SQ_THROW_IF(nCols <= 0 || nCols % 16 != 0, "field GEMM: %d output columns is not a positive multiple of 16",
                nCols);
    SQ_THROW_IF(K <= 0 || K % 256 != 0 || Nw * 32 != K, "field GEMM: K=%d must be a positive multiple of 256", K);
    SQ_THROW_IF(ld < K, "field GEMM: row stride %d is shorter than K=%d", ld, K);
}

}  // namespace

size_t fieldFragElemsBF16x2(int nCols, int K) { return 2 * (size_t)nCols * K; } /* hi and lo planes, uint16 each */
size_t fieldFragElemsF32(int nCols, int K) { return 3 * (size_t)nCols * K; } /* hi, mid and lo planes */
size_t fieldFragElemsFP8(int nCols, int K) { return (size_t)nCols * K; } /* bytes */

void buildFragFP8(const uint8_t *d_B8, int rows, int ld8, int nCols, int K, uint8_t *d_frag, size_t fragElems,
                  hipStream_t stream) {
    checkGemm(nCols, K, ld8, K / 32);
    SQ_THROW_IF(ld8 % 16 != 0, "fp8 fragments: row stride %d is not a multiple of 16 bytes", ld8);
    SQ_THROW_IF(rows > nCols, "field GEMM: %d rows exceed %d columns", rows, nCols);
    SQ_THROW_IF(fragElems < fieldFragElemsFP8(nCols, K), "fp8 fragment buffer holds %zu of %zu bytes", fragElems,
                fieldFragElemsFP8(nCols, K));
    const size_t threads = (size_t)(nCols / 16) * (K / 128) * kWave;
    buildFragFP8Kernel<<<(unsigned)((threads + 255) / 256), 256, 0, stream>>>(d_B8, rows, ld8, nCols / 16, K,
                                                                              reinterpret_cast<uint4 *>(d_frag));
    HIP_CHECK(hipGetLastError());
}

void fieldRefreshGrouped(int fieldFormat, const FieldGemmJob *d_jobs, const int *d_blockStart, int nJobs,
                         int totalBlocks, hipStream_t stream) {
    SQ_THROW_IF(nJobs <= 0 || totalBlocks <= 0, "grouped field GEMM: no work");
    if (fieldFormat == fieldBF16x2)
        fieldRefreshGroupedKernel<fieldBF16x2><<<totalBlocks, kBlockThreads, 0, stream>>>(d_jobs, d_blockStart, nJobs);
    else if (fieldFormat == fieldFP8)
        fieldRefreshGroupedKernel<fieldFP8><<<totalBlocks, kBlockThreads, 0, stream>>>(d_jobs, d_blockStart, nJobs);
    else
        fieldRefreshGroupedKernel<fieldF32><<<totalBlocks, kBlockThreads, 0, stream>>>(d_jobs, d_blockStart, nJobs);
    HIP_CHECK(hipGetLastError());
}

int fieldGemmJobBlocks(const FieldGemmJob &j, int fieldFormat) {
    return j.nCols / (fieldFormat != fieldFP8 && bf16ColSplit(j.K) ? 16 * kSplitK : 16);
}

void checkFieldGemmJob(const FieldGemmJob &j, int fieldFormat) {
    checkGemm(j.nCols, j.K, j.K, j.Nw);
    SQ_THROW_IF(j.ldF < j.nCols || !j.frag || !j.Qbits || !j.F || (fieldFormat == fieldFP8 && !j.colScale),
                "grouped field GEMM: bad job");
}

void fieldRefreshFP8(const uint8_t *d_frag, size_t fragElems, const float *d_colScale, const uint32_t *d_Qbits,
                     int nCols, int K, int m, int Nw, int ldF, float *d_F, hipStream_t stream) {
    checkGemm(nCols, K, K, Nw);
    SQ_THROW_IF(ldF < nCols, "field GEMM: output stride %d < %d columns", ldF, nCols);
    SQ_THROW_IF(fragElems < fieldFragElemsFP8(nCols, K), "fp8 fragment buffer too small");
    SQ_THROW_IF(!d_colScale, "fp8 field GEMM: no column scales");
    fieldRefreshFP8Kernel<<<nCols / 16, kBlockThreads, 0, stream>>>(reinterpret_cast<const uint4 *>(d_frag),
                                                                    d_colScale, d_Qbits, K, m, Nw, ldF, d_F);
    HIP_CHECK(hipGetLastError());
}

void buildFragBF16x2(const float *d_B, int rows, int ld, int nCols, int K, uint16_t *d_frag, size_t fragElems,
                     hipStream_t stream) {
    checkGemm(nCols, K, ld, K / 32);
    SQ_THROW_IF(rows > nCols, "field GEMM: %d rows exceed %d columns", rows, nCols);
    SQ_THROW_IF(fragElems < fieldFragElemsBF16x2(nCols, K), "bf16x2 fragment buffer holds %zu of %zu elements",
                fragElems, fieldFragElemsBF16x2(nCols, K));
    const size_t threads = (size_t)(nCols / 16) * (K / 32) * kWave;
    buildFragBF16Kernel<2><<<(unsigned)((threads + 255) / 256), 256, 0, stream>>>(d_B, rows, ld, nCols / 16, K,
                                                                               reinterpret_cast<uint4 *>(d_frag));
    HIP_CHECK(hipGetLastError());
}

void buildFragF32(const float *d_B, int rows, int ld, int nCols, int K, uint16_t *d_frag, size_t fragElems,
                  hipStream_t stream) {
    checkGemm(nCols, K, ld, K / 32);
    SQ_THROW_IF(rows > nCols, "field GEMM: %d rows exceed %d columns", rows, nCols);
    SQ_THROW_IF(fragElems < fieldFragElemsF32(nCols, K), "fp32 fragment buffer holds %zu of %zu elements", fragElems,
                fieldFragElemsF32(nCols, K));
    const size_t threads = (size_t)(nCols / 16) * (K / 32) * kWave;
    buildFragBF16Kernel<3><<<(unsigned)((threads + 255) / 256), 256, 0, stream>>>(d_B, rows, ld, nCols / 16, K,
                                                                               reinterpret_cast<uint4 *>(d_frag));
    HIP_CHECK(hipGetLastError());
}

void fieldRefreshBF16x2(const uint16_t *d_frag, size_t fragElems, const uint32_t *d_Qbits, int nCols, int K, int m,
                        int Nw, int ldF, float *d_F, hipStream_t stream) {
    checkGemm(nCols, K, K, Nw);
    SQ_THROW_IF(ldF < nCols, "field GEMM: output stride %d < %d columns", ldF, nCols);
    SQ_THROW_IF(fragElems < fieldFragElemsBF16x2(nCols, K), "bf16x2 fragment buffer too small");
    fieldRefreshBF16x2Kernel<<<nCols / 16, kBlockThreads, 0, stream>>>(reinterpret_cast<const uint4 *>(d_frag),
                                                                       d_Qbits, K, m, Nw, ldF, d_F);
    HIP_CHECK(hipGetLastError());
}

void fieldRefreshF32MFMA(const uint16_t *d_frag, size_t fragElems, const uint32_t *d_Qbits, int nCols, int K,
                         int m, int Nw, int ldF, float *d_F, hipStream_t stream) {
    checkGemm(nCols, K, K, Nw);
    SQ_THROW_IF(ldF < nCols, "field GEMM: output stride %d < %d columns", ldF, nCols);
    SQ_THROW_IF(fragElems < fieldFragElemsF32(nCols, K), "fp32 fragment buffer too small");
    fieldRefreshBF16x3Kernel<<<nCols / 16, kBlockThreads, 0, stream>>>(reinterpret_cast<const uint4 *>(d_frag),
                                                                       d_Qbits, K, m, Nw, ldF, d_F);
    HIP_CHECK(hipGetLastError());
}

}  // namespace sqaod_rocm
