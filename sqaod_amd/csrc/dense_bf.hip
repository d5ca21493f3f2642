/*
 * Dense brute-force search: E(x) = x^T W x over every x in [xb, xe), keeping
 * the minimum and the ascending list of the x that reach it.  The semantics
 * are those of one sqaod.cpu batch searcher (CPUDenseGraphBatchSearch.cpp:
 * 26-50) and of the CUDA device's tile merge (CUDADenseGraphBFSearcher.cpp:
 * 155-185): a strictly lower minimum replaces the list, an equal one appends,
 * and the list keeps the first `cap` (= tile size) x in ascending order.
 *
 * Bits: bit (N-1-pos) of the packed x is x[pos] (Common.cpp:79-93), so the L
 * low bits of x are the last L positions.  Split x = (xH, xL) with xL the low
 * L = min(N, 12) bits:
 *   E(x) = EH(xH) + EL(xL) + sum_{b in xL} c_b(xH),
 *   EH   = xH-part of x^T W x,   EL = xL-part (a 2^L table, computed once),
 *   c_b  = 2 sum_{pos in H} W[N-1-b][pos] x_pos   (W symmetric),
 * and the cross sum is read from two 64-entry subset-sum tables (bits 0..5
 * and 6..11 of xL).  A candidate then costs three LDS reads and three double
 * adds; the energy is formed in double and rounded to the solver's real once,
 * and the same function is used by every kernel, so the scan and the select
 * agree bit for bit.
 *
 * Work split: one block owns a unit of R consecutive high patterns times all
 * 2^L low patterns (R = min(16, 2^(N-L)), up to 65,536 candidates).
 *   bfScan   : per block, min E and the number of x reaching it;
 *   bfMerge  : one block: the range minimum, the list update rule, and each
 *              block's output offset (exclusive scan over blocks);
 *   bfSelect : blocks holding the range minimum write their x in ascending
 *              order at their offset (ordered compaction), up to cap.
 * Nothing is read back to the host between ranges.
 */
#include <cfloat>

#include "common.h"
#include "kernels.h"

namespace sqaod_rocm {

namespace {

constexpr int kThreads = 256;
constexpr int kWave = 64;
constexpr unsigned long long kNoOffset = ~0ull;

/* LDS tables of one block */
struct BFTables {
    double EL[1 << kBFMaxLowBits];
    double P0[kBFMaxHighPatterns][64];
    double P1[kBFMaxHighPatterns][64];
    double C[kBFMaxHighPatterns][kBFMaxLowBits];
    double row[kBFMaxHighPatterns][64];
    double EH[kBFMaxHighPatterns];
};

/* x_pos of the high pattern xH (positions 0 .. nH-1) */
__device__ __forceinline__ int highBit(unsigned long long xH, int nH, int pos) {
    return (int)((xH >> (nH - 1 - pos)) & 1ull);
}

/* Builds the tables of the block's R high patterns xH0 .. xH0+R-1. */
__device__ void buildTables(BFTables &t, const double *W, const double *EL, const BFPlan &p,
                            unsigned long long xH0) {
    const int N = p.N, L = p.L, nH = p.nH, R = p.R;
    for (int i = threadIdx.x; i < (1 << L); i += kThreads)
        t.EL[i] = EL[i];
    /* c_b(xH) for b < L and the row sums x_pos * sum_{pos'} W[pos][pos'] x_pos' */
    for (int task = threadIdx.x; task < R * N; task += kThreads) {
        const int r = task / N, k = task % N;
        const unsigned long long xH = xH0 + (unsigned long long)r;
        if (k < L) {
            const double *Wr = W + (size_t)(N - 1 - k) * N;
            double c = 0.;
            for (int pos = 0; pos < nH; ++pos)
                if (highBit(xH, nH, pos))
                    c += Wr[pos];
            t.C[r][k] = 2. * c;
        } else {
            const int pos = k - L;
            double v = 0.;
            if (highBit(xH, nH, pos)) {
                const double *Wr = W + (size_t)pos * N;
                for (int q = 0; q < nH; ++q)
                    if (highBit(xH, nH, q))
                        v += Wr[q];
            }
            t.row[r][pos] = v;
        }
    }
    __syncthreads();
    /* subset sums of c over bits 0..Llo-1 and Llo..L-1 of xL, and EH */
    const int Llo = p.Llo, Lhi = L - p.Llo;
    for (int task = threadIdx.x; task < R * 128; task += kThreads) {
        const int r = task / 128, a = task % 128;
        double s = 0.;
        if (a < 64) {
            for (int b = 0; b < Llo; ++b)
                if ((a >> b) & 1)
                    s += t.C[r][b];
            t.P0[r][a] = s;
        } else {
            const int aa = a - 64;
            for (int b = 0; b < Lhi; ++b)
                if ((aa >> b) & 1)
                    s += t.C[r][Llo + b];
            t.P1[r][aa] = s;
        }
    }
    if (threadIdx.x < (unsigned)R) {
        double s = 0.;
        for (int pos = 0; pos < nH; ++pos)
            s += t.row[threadIdx.x][pos];
        t.EH[threadIdx.x] = s;
    }
    __syncthreads();
}

template <class real>
__device__ __forceinline__ real energyAt(const BFTables &t, const BFPlan &p, int r, int xL) {
    const double e = ((t.EH[r] + t.EL[xL]) + t.P0[r][xL & ((1 << p.Llo) - 1)]) + t.P1[r][xL >> p.Llo];
    return (real)e;
}

template <class real>
__device__ __forceinline__ real waveMin(real v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const real w = __shfl_xor(v, o, kWave);
        v = w < v ? w : v;
    }
    return v;
}

__device__ __forceinline__ unsigned waveSum(unsigned v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    return v;
}

template <class real>
__global__ __launch_bounds__(kThreads) void bfScanKernel(const double *W, const double *EL, BFPlan p,
                                                         unsigned long long xb, unsigned long long xe,
                                                         unsigned long long unit0, real *blockMin,
                                                         unsigned *blockCnt) {
    __shared__ BFTables t;
    __shared__ real wMin[kThreads / kWave];
    __shared__ unsigned wCnt[kThreads / kWave];
    const unsigned long long unit = unit0 + blockIdx.x;
    const unsigned long long xH0 = unit * (unsigned long long)p.R;
    buildTables(t, W, EL, p, xH0);

    const real inf = (real)HUGE_VAL;
    real best = inf;
    unsigned cnt = 0;
    for (int xL = threadIdx.x; xL < (1 << p.L); xL += kThreads) {
        for (int r = 0; r < p.R; ++r) {
            const unsigned long long x = ((xH0 + (unsigned long long)r) << p.L) | (unsigned long long)xL;
            if (x < xb || x >= xe)
                continue;
            const real e = energyAt<real>(t, p, r, xL);
            if (e < best) {
                best = e;
                cnt = 1;
            } else if (e == best) {
                ++cnt;
            }
        }
    }
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    real m = waveMin(best);
    if (lane == 0)
        wMin[w] = m;
    __syncthreads();
    m = wMin[0];
    for (int i = 1; i < kThreads / kWave; ++i)
        m = wMin[i] < m ? wMin[i] : m;
    const unsigned c = waveSum(best == m ? cnt : 0u);
    if (lane == 0)
        wCnt[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned s = 0;
        for (int i = 0; i < kThreads / kWave; ++i)
            s += wCnt[i];
        blockMin[blockIdx.x] = m;
        blockCnt[blockIdx.x] = s;
    }
}

/* One block.  Applies the list rule of CPUDenseGraphBatchSearch.cpp:33-48 to
 * the range as a whole and assigns each block holding the range minimum its
 * output offset (kNoOffset: nothing to write). */
template <class real>
__global__ __launch_bounds__(kThreads) void bfMergeKernel(const real *blockMin, const unsigned *blockCnt, int nb,
                                                          BFState<real> *st, unsigned long long cap,
                                                          unsigned long long *offsets) {
    __shared__ real wMin[kThreads / kWave];
    __shared__ unsigned long long wSum[kThreads / kWave];
    __shared__ real sMin;
    __shared__ unsigned long long sBase;
    __shared__ int sActive;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    real m = (real)HUGE_VAL;
    for (int b = threadIdx.x; b < nb; b += kThreads)
        m = blockMin[b] < m ? blockMin[b] : m;
    m = waveMin(m);
    if (lane == 0)
        wMin[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        real tm = wMin[0];
        for (int i = 1; i < kThreads / kWave; ++i)
            tm = wMin[i] < tm ? wMin[i] : tm;
        const real Emin = st->Emin;
        int active = 0;
        unsigned long long base = 0;
        if (tm < Emin) { /* a lower minimum: the list restarts */
            st->Emin = tm;
            base = 0;
            active = 1;
        } else if (tm == Emin) { /* ties append while the list is short of cap */
            base = st->count;
            active = base < cap;
        }
        sMin = tm;
        sBase = base;
        sActive = active;
    }
    __syncthreads();
    const real tm = sMin;
    const int active = sActive;
    unsigned long long running = sBase;
    /* exclusive scan over blocks, 256 at a time, in block order */
    for (int b0 = 0; b0 < nb; b0 += kThreads) {
        const int b = b0 + threadIdx.x;
        const bool hit = active && b < nb && blockMin[b] == tm;
        const unsigned long long v = hit ? blockCnt[b] : 0ull;
        /* inclusive wave scan */
        unsigned long long incl = v;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const unsigned long long u = __shfl_up(incl, o, kWave);
            if (lane >= o)
                incl += u;
        }
        if (lane == kWave - 1)
            wSum[w] = incl;
        __syncthreads();
        unsigned long long before = running;
        for (int i = 0; i < w; ++i)
            before += wSum[i];
        const unsigned long long off = before + incl - v;
        if (b < nb)
            offsets[b] = (hit && off < cap) ? off : kNoOffset;
        unsigned long long total = 0;
        for (int i = 0; i < kThreads / kWave; ++i)
            total += wSum[i];
        running += total;
        __syncthreads();
    }
    if (threadIdx.x == 0 && active)
        st->count = running < cap ? running : cap;
}

/* Blocks with an offset write their x with E == Emin in ascending order. */
template <class real>
__global__ __launch_bounds__(kThreads) void bfSelectKernel(const double *W, const double *EL, BFPlan p,
                                                           unsigned long long xb, unsigned long long xe,
                                                           unsigned long long unit0, const unsigned long long *offsets,
                                                           const BFState<real> *st, unsigned long long cap,
                                                           unsigned long long *xList) {
    const unsigned long long off0 = offsets[blockIdx.x];
    if (off0 == kNoOffset) /* block-uniform */
        return;
    __shared__ BFTables t;
    __shared__ unsigned long long wSum[kThreads / kWave];
    const unsigned long long unit = unit0 + blockIdx.x;
    const unsigned long long xH0 = unit * (unsigned long long)p.R;
    buildTables(t, W, EL, p, xH0);
    const real Emin = st->Emin;

    /* thread t owns the contiguous index run [t*per, (t+1)*per) of the
     * block's candidates i = r * 2^L + xL, i.e. ascending x */
    const int U = p.R << p.L;
    const int per = (U + kThreads - 1) / kThreads;
    const int i0 = threadIdx.x * per;
    const int i1 = (i0 + per) < U ? (i0 + per) : U;
    auto tie = [&](int i, unsigned long long &x) {
        const int r = i >> p.L, xL = i & ((1 << p.L) - 1);
        x = ((xH0 + (unsigned long long)r) << p.L) | (unsigned long long)xL;
        return x >= xb && x < xe && energyAt<real>(t, p, r, xL) == Emin;
    };
    unsigned long long cnt = 0, x;
    for (int i = i0; i < i1; ++i)
        cnt += tie(i, x) ? 1u : 0u;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    unsigned long long incl = cnt;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const unsigned long long u = __shfl_up(incl, o, kWave);
        if (lane >= o)
            incl += u;
    }
    if (lane == kWave - 1)
        wSum[w] = incl;
    __syncthreads();
    unsigned long long pos = off0 + incl - cnt;
    for (int i = 0; i < w; ++i)
        pos += wSum[i];
    for (int i = i0; i < i1 && pos < cap; ++i)
        if (tie(i, x))
            xList[pos++] = x;
}

}  // namespace

BFPlan bfPlan(int N) {
    BFPlan p;
    p.N = N;
    p.L = N < kBFMaxLowBits ? N : kBFMaxLowBits;
    p.nH = N - p.L;
    p.Llo = p.L < 6 ? p.L : 6;
    p.lgR = p.nH < 4 ? p.nH : 4;
    p.R = 1 << p.lgR; /* at most kBFMaxHighPatterns */
    return p;
}

template <class real>
void bfSearchRange(const double *d_W, const double *d_EL, const BFPlan &p, unsigned long long xb,
                   unsigned long long xe, BFState<real> *d_state, unsigned long long cap, real *d_blockMin,
                   unsigned *d_blockCnt, unsigned long long *d_offsets, int maxBlocks,
                   unsigned long long *d_xList, hipStream_t stream) {
    if (xb >= xe)
        return;
    const int lg = p.L + p.lgR; /* log2 of the candidates per block */
    const unsigned long long unit0 = xb >> lg;
    const unsigned long long unit1 = (xe + (1ull << lg) - 1) >> lg;
    const unsigned long long nb = unit1 - unit0;
    SQ_THROW_IF(nb > (unsigned long long)maxBlocks, "bf search range of %llu units exceeds %d", nb, maxBlocks);
    bfScanKernel<real><<<(unsigned)nb, kThreads, 0, stream>>>(d_W, d_EL, p, xb, xe, unit0, d_blockMin, d_blockCnt);
    HIP_CHECK(hipGetLastError());
    bfMergeKernel<real><<<1, kThreads, 0, stream>>>(d_blockMin, d_blockCnt, (int)nb, d_state, cap, d_offsets);
    HIP_CHECK(hipGetLastError());
    bfSelectKernel<real><<<(unsigned)nb, kThreads, 0, stream>>>(d_W, d_EL, p, xb, xe, unit0, d_offsets, d_state, cap,
                                                                d_xList);
    HIP_CHECK(hipGetLastError());
}

template void bfSearchRange<float>(const double *, const double *, const BFPlan &, unsigned long long,
                                   unsigned long long, BFState<float> *, unsigned long long, float *, unsigned *,
                                   unsigned long long *, int, unsigned long long *, hipStream_t);
template void bfSearchRange<double>(const double *, const double *, const BFPlan &, unsigned long long,
                                    unsigned long long, BFState<double> *, unsigned long long, double *, unsigned *,
                                    unsigned long long *, int, unsigned long long *, hipStream_t);

}  // namespace sqaod_rocm
