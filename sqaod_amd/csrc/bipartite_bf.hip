/*
 * Bipartite brute-force search: E(x0, x1) = b0.x0 + b1.x1 + x1^T W x0 over a
 * rectangle of (x0, x1) pairs (calculate_E_2d, SharedFormulas.cpp:123-144),
 * keeping the minimum and the pairs that reach it.  The list semantics are
 * one sqaod.cpu worker's (CPUBipartiteGraphBatchSearch.cpp:29-62 driven by
 * searchRangeSingleThread, CPUBipartiteGraphBFSearcher.cpp:166-198): pairs
 * are met in the order (x0 / T0, x1, x0 % T0) with T0 the x0 tile size, a
 * strictly lower E restarts the list and an equal E appends while the list
 * holds fewer than N0 + N1 pairs -- so the list is the N0 + N1 smallest
 * visiting keys among the minimisers.
 *
 * For a fixed x1, E is linear in x0:  E = B(x1) + sum_{pos in x0} d_pos(x1)
 * with d_pos = b0[pos] + sum_{i in x1} W[i][pos] and B = b1.x1.  Split x0 =
 * (xH, xL) with xL its low L = min(N0, 10) bits (bit N0-1-pos of a packed x
 * is x[pos], Common.cpp:79-86): E = (B + SH(xH)) + TL(xL), both subset sums
 * of d, formed in double and rounded to the solver's real once.
 *
 * Block = (one x1, R = 16 consecutive xH), 256 threads, each thread owning
 * the low patterns xL = t, t + 256, ...; a candidate costs one double add and
 * a compare.
 *   bgScan   : per block, the minimum of its candidates;
 *   bgSelect : a block whose minimum equals the range minimum writes its
 *              first K minimisers in ascending x0 (for a fixed x1 that is the
 *              visiting order) and their count.
 * The host merges the blocks' lists by visiting key.
 */
#include <cfloat>

#include "common.h"
#include "kernels.h"

namespace sqaod_rocm {

namespace {

constexpr int kThreads = 256;
constexpr int kLowMax = 10;

struct BGBlockTables {
    double d[64];
    double SH[kBGBFHighPerBlock];
    double B;
};

/* the block's x1 and first xH from its index */
__device__ __forceinline__ void blockCoords(const BGBFArgs &a, unsigned long long &x1, unsigned long long &xH0) {
    const unsigned long long b = blockIdx.x;
    x1 = a.x1b + b % a.nX1;
    xH0 = a.hb + (b / a.nX1) * kBGBFHighPerBlock;
}

template <class real>
__device__ void buildBlock(const BGBFArgs &a, const real *b0, const real *b1, const real *W, BGBlockTables &t,
                           unsigned long long x1, unsigned long long xH0) {
    const int N0 = a.N0, N1 = a.N1;
    if ((int)threadIdx.x < N0) {
        const int pos = threadIdx.x;
        double s = a.sgn * (double)b0[pos];
        for (int i = 0; i < N1; ++i)
            if ((x1 >> (N1 - 1 - i)) & 1ull)
                s += a.sgn * (double)W[(size_t)i * N0 + pos];
        t.d[pos] = s;
    }
    if (threadIdx.x == 0) {
        double s = 0.;
        for (int i = 0; i < N1; ++i)
            if ((x1 >> (N1 - 1 - i)) & 1ull)
                s += a.sgn * (double)b1[i];
        t.B = s;
    }
    __syncthreads();
    if ((int)threadIdx.x < kBGBFHighPerBlock) {
        const unsigned long long xH = xH0 + threadIdx.x;
        double s = 0.;
        for (int b = 0; b < a.nH; ++b)
            if ((xH >> b) & 1ull)
                s += t.d[N0 - 1 - (a.L + b)];
        t.SH[threadIdx.x] = s;
    }
    __syncthreads();
}

/* TL of the low pattern xL */
__device__ __forceinline__ double lowSum(const BGBFArgs &a, const BGBlockTables &t, int xL) {
    double s = 0.;
    for (int b = 0; b < a.L; ++b)
        if ((xL >> b) & 1)
            s += t.d[a.N0 - 1 - b];
    return s;
}

template <class real>
__device__ __forceinline__ real pairEnergy(const BGBlockTables &t, int r, double tl) {
    return (real)((t.B + t.SH[r]) + tl);
}

template <class real>
__global__ __launch_bounds__(kThreads) void bgScanKernel(BGBFArgs a, const real *b0, const real *b1, const real *W,
                                                         real *blockMin) {
    __shared__ BGBlockTables t;
    __shared__ real red[kThreads / 64];
    unsigned long long x1, xH0;
    blockCoords(a, x1, xH0);
    buildBlock<real>(a, b0, b1, W, t, x1, xH0);
    const int nL = 1 << a.L;
    const int nR = (int)((a.he - xH0) < (unsigned long long)kBGBFHighPerBlock ? (a.he - xH0) : kBGBFHighPerBlock);
    real mn = (real)FLT_MAX;
    for (int xL = threadIdx.x; xL < nL; xL += kThreads) {
        const double tl = lowSum(a, t, xL);
        for (int r = 0; r < nR; ++r) {
            const real e = pairEnergy<real>(t, r, tl);
            mn = e < mn ? e : mn;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const real o = __shfl_xor(mn, off);
        mn = o < mn ? o : mn;
    }
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x / 64] = mn;
    __syncthreads();
    if (threadIdx.x == 0) {
        real m = red[0];
        for (int w = 1; w < kThreads / 64; ++w)
            m = red[w] < m ? red[w] : m;
        blockMin[blockIdx.x] = m;
    }
}

template <class real>
__global__ __launch_bounds__(kThreads) void bgSelectKernel(BGBFArgs a, const real *b0, const real *b1, const real *W,
                                                           const real *blockMin, real Emin,
                                                           unsigned long long *pairs, int *counts) {
    if (blockMin[blockIdx.x] != Emin) {
        if (threadIdx.x == 0)
            counts[blockIdx.x] = 0;
        return;
    }
    __shared__ BGBlockTables t;
    __shared__ int wsum[kThreads / 64];
    __shared__ int base;
    unsigned long long x1, xH0;
    blockCoords(a, x1, xH0);
    buildBlock<real>(a, b0, b1, W, t, x1, xH0);
    if (threadIdx.x == 0)
        base = 0;
    __syncthreads();
    const int nL = 1 << a.L;
    const int nR = (int)((a.he - xH0) < (unsigned long long)kBGBFHighPerBlock ? (a.he - xH0) : kBGBFHighPerBlock);
    const int lane = threadIdx.x & 63, w = threadIdx.x / 64;
    unsigned long long *out = pairs + (size_t)blockIdx.x * a.K * 2;
    /* ascending x0 = (r, xL): r outer, then xL = k * 256 + t (k outer, t inner) */
    for (int r = 0; r < nR; ++r) {
        for (int k0 = 0; k0 < nL; k0 += kThreads) {
            const int xL = k0 + (int)threadIdx.x;
            bool tie = false;
            if (xL < nL)
                tie = pairEnergy<real>(t, r, lowSum(a, t, xL)) == Emin;
            const unsigned long long bal = __ballot(tie);
            if (lane == 0)
                wsum[w] = __popcll(bal);
            __syncthreads();
            int before = base;
            for (int v = 0; v < w; ++v)
                before += wsum[v];
            const int pos = before + __popcll(bal & ((1ull << lane) - 1ull));
            if (tie && pos < a.K) {
                out[2 * pos] = ((xH0 + (unsigned long long)r) << a.L) | (unsigned long long)xL;
                out[2 * pos + 1] = x1;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int tot = 0;
                for (int v = 0; v < kThreads / 64; ++v)
                    tot += wsum[v];
                base += tot;
            }
            __syncthreads();
            if (base >= a.K)
                break;
        }
        if (base >= a.K)
            break;
    }
    if (threadIdx.x == 0)
        counts[blockIdx.x] = base < a.K ? base : a.K;
}

}  // namespace

BGBFArgs bgBFPlan(int N0, int N1) {
    BGBFArgs a{};
    a.N0 = N0;
    a.N1 = N1;
    a.L = N0 < kLowMax ? N0 : kLowMax;
    a.nH = N0 - a.L;
    a.K = N0 + N1;
    return a;
}

template <class real>
void bgBFScan(const BGBFArgs &a, const real *d_b0, const real *d_b1, const real *d_W, real *d_blockMin,
              hipStream_t stream) {
    SQ_THROW_IF(a.N0 < 1 || a.N0 > 63 || a.N1 < 1 || a.N1 > 63 || a.L != (a.N0 < kLowMax ? a.N0 : kLowMax),
                "bipartite BF: bad plan");
    SQ_THROW_IF(a.he <= a.hb || a.nX1 == 0 || a.x1b + a.nX1 > (1ull << a.N1) || a.he > (1ull << a.nH),
                "bipartite BF: bad range");
    bgScanKernel<real><<<(unsigned)bgBFBlocks(a), kThreads, 0, stream>>>(a, d_b0, d_b1, d_W, d_blockMin);
    HIP_CHECK(hipGetLastError());
}

template <class real>
void bgBFSelect(const BGBFArgs &a, const real *d_b0, const real *d_b1, const real *d_W, const real *d_blockMin,
                real Emin, unsigned long long *d_pairs, int *d_counts, hipStream_t stream) {
    bgSelectKernel<real><<<(unsigned)bgBFBlocks(a), kThreads, 0, stream>>>(a, d_b0, d_b1, d_W, d_blockMin, Emin,
                                                                           d_pairs, d_counts);
    HIP_CHECK(hipGetLastError());
}

unsigned long long bgBFBlocks(const BGBFArgs &a) {
    return a.nX1 * ((a.he - a.hb + kBGBFHighPerBlock - 1) / kBGBFHighPerBlock);
}

template void bgBFScan<float>(const BGBFArgs &, const float *, const float *, const float *, float *, hipStream_t);
template void bgBFScan<double>(const BGBFArgs &, const double *, const double *, const double *, double *,
                               hipStream_t);
template void bgBFSelect<float>(const BGBFArgs &, const float *, const float *, const float *, const float *, float,
                                unsigned long long *, int *, hipStream_t);
template void bgBFSelect<double>(const BGBFArgs &, const double *, const double *, const double *, const double *,
                                 double, unsigned long long *, int *, hipStream_t);

}  // namespace sqaod_rocm
