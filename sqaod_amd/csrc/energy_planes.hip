/*
 * Annealer energies on the integer MFMA units (gfx950), exact or to FP64
 * accuracy:
 *   dense     E_y = -c - sum_i q_yi (h_i + sum_j J_ij q_yj)
 *             (DGFuncs::calculate_E, sqaodc/cpu/SharedFormulas.cpp:69-83);
 *   bipartite E_y = -c - h0.q0_y - sum_i q1_yi (h1_i + sum_j J_ij q0_yj)
 *             (BGFuncs::calculate_E, SharedFormulas.cpp:184-202).
 * The reference forms the same quantity with a GEMM (the CUDA device's
 * DeviceFormulas.cpp:42-49 -> DeviceMath.cpp:192-209) in `real`.
 *
 * Fixed-point digit planes.  The matrix B (nOut x K: J, or the bipartite J
 * with rows for q1 and columns for q0) is written once per problem as
 *     B_ik = 2^es * V_ik,   V_ik = sum_{p<P} 256^p D_p,ik,   D in [-128, 127],
 * with es the exponent of the smallest set bit over all elements, so that
 * every V is an integer and the representation is exact, whenever that takes
 * at most Pmax planes (6 for float, 7 for double).  Otherwise es is raised to
 * keep P = Pmax and elements are rounded to the nearest multiple of 2^es,
 * i.e. to 8 Pmax - 2 bits below the largest exponent (2^-46 relative to
 * max|B| for float, 2^-54 for double).  Integer problems take one or two
 * planes.
 *
 * Per plane the products D_p . q (q = +-1) are int8 x int8 with exact int32
 * accumulation (|sum| <= 128 K), on v_mfma_i32_16x16x64_i8: A = a 16-row
 * tile of the plane (laid out in fragment order, one 1 KiB coalesced load per
 * wave), B = 16 slices' spins expanded from their packed bits in registers.
 * The planes are combined per element in double by Horner's rule (exact while
 * the field is below 2^53 2^es), multiplied by q_yi (plus h_i in the first
 * k-chunk) and reduced over the wave's 16 rows in a fixed order; partial sums
 * per (row tile, k-chunk) go to memory and one workgroup per slice adds them
 * in a fixed order.  The result is deterministic, exact on integer problems,
 * and for real data differs from an FP64 evaluation only by FP64 rounding of
 * the final sums.
 */
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <climits>
#include <cmath>

namespace sqaod_rocm {

namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kScanThreads = 256;
constexpr int kFinalThreads = 256;

/* exponents of the highest and lowest set bits of |x| (x finite, non-zero) */
__device__ __forceinline__ void bitExponents(float x, int &hi, int &lo) {
    const uint32_t u = __float_as_uint(x) & 0x7FFFFFFFu;
    const int e = (int)(u >> 23);
    const uint32_t sig = e ? ((u & 0x7FFFFFu) | 0x800000u) : u;
    const int e0 = e ? e - 150 : -149;
    lo = e0 + __builtin_ctz(sig);
    hi = e0 + 31 - __builtin_clz(sig);
}
__device__ __forceinline__ void bitExponents(double x, int &hi, int &lo) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x) & 0x7FFFFFFFFFFFFFFFull;
    const int e = (int)(u >> 52);
    const unsigned long long sig = e ? ((u & 0xFFFFFFFFFFFFFull) | 0x10000000000000ull) : u;
    const int e0 = e ? e - 1075 : -1074;
    lo = e0 + __builtin_ctzll(sig);
    hi = e0 + 63 - __builtin_clzll(sig);
}

/* scan[0] = max hi, scan[1] = min lo over non-zero elements, scan[2] = 1 if
 * some element is not finite */
template <class real>
__global__ __launch_bounds__(kScanThreads) void planeScanKernel(const real *B, long long sRow, long long sCol, int nOut,
                                                                int K, int *scan) {
    int hiMax = INT_MIN, loMin = INT_MAX, bad = 0;
    for (int i = blockIdx.x; i < nOut; i += gridDim.x) {
        const real *row = B + (long long)i * sRow;
        for (int k = threadIdx.x; k < K; k += kScanThreads) {
            const real x = row[(long long)k * sCol];
            if (!isfinite(x)) {
                bad = 1;
            } else if (x != (real)0) {
                int hi, lo;
                bitExponents(x, hi, lo);
                hiMax = hi > hiMax ? hi : hiMax;
                loMin = lo < loMin ? lo : loMin;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const int h = __shfl_xor(hiMax, off), l = __shfl_xor(loMin, off);
        hiMax = h > hiMax ? h : hiMax;
        loMin = l < loMin ? l : loMin;
        bad |= __shfl_xor(bad, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (hiMax != INT_MIN)
            atomicMax(&scan[0], hiMax);
        if (loMin != INT_MAX)
            atomicMin(&scan[1], loMin);
        if (bad)
            atomicOr(&scan[2], 1);
    }
}

/* One thread per (row tile, k-block, lane): the lane's 16 elements
 * B(16 it + (lane & 15), 64 kb + 16 (lane >> 4) + b), b < 16, as P digit
 * bytes, one uint4 per plane. */
template <class real>
__global__ __launch_bounds__(256) void planeBuildKernel(const real *B, long long sRow, long long sCol, int nOut, int K,
                                                        int outTiles, int kBlocks, int P, int es, uint4 *frag) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long perPlane = (long long)outTiles * kBlocks * kWave;
    if (t >= perPlane)
        return;
    const int lane = (int)(t & 63);
    const long long rest = t >> 6;
    const int kb = (int)(rest % kBlocks), it = (int)(rest / kBlocks);
    const int i = 16 * it + (lane & 15);
    const int k0 = 64 * kb + 16 * (lane >> 4);
    uint32_t w[8][4]; /* [plane][dword] */
#pragma unroll
    for (int p = 0; p < 8; ++p)
#pragma unroll
        for (int d = 0; d < 4; ++d)
            w[p][d] = 0u;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        const int k = k0 + b;
        long long V = 0;
        if (i < nOut && k < K) {
            const double x = (double)B[(long long)i * sRow + (long long)k * sCol];
            V = (long long)rint(ldexp(x, -es));
        }
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            if (p < P) {
                const int d = ((int)(V & 255) ^ 128) - 128; /* balanced digit */
                V = (V - d) >> 8;
                w[p][b >> 2] |= ((uint32_t)d & 255u) << (8 * (b & 3));
            }
        }
    }
#pragma unroll
    for (int p = 0; p < 8; ++p)
        if (p < P)
            frag[p * perPlane + t] = make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
}

/* 16 spin bits (bit b = site k0 + b, 1 = +1) as 16 int8 +-1: each nibble is
 * spread to one bit per byte, then v_perm picks 0x01 or 0xFF per byte. */
__device__ __forceinline__ i32x4 spinsI8(uint32_t bits) {
    i32x4 v;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const uint32_t nib = (bits >> (4 * d)) & 15u;
        const uint32_t sel = (nib * 0x00204081u) & 0x01010101u;
        v[d] = (int)__builtin_amdgcn_perm(0x000001FFu, 0x000001FFu, sel);
    }
    return v;
}

template <class real>
struct PlaneArgs {
    const uint4 *frag;
    const real *hOut;
    const uint32_t *Qin, *Qout;
    double *partial;
    double scale;
    int nOut, outTiles, kBlocks, NwIn, NwOut, m;
    int nYG, KS, chunk, nWaves, nSlots;
};

/* wave wi = (kc * outTiles + it) * nYG + yg: rows [16 it, +16), slices
 * [16 TY yg, +16 TY), k-blocks [chunk kc, +chunk) */
template <class real, int P, int TY>
__global__ __launch_bounds__(kWavesPerBlock *kWave) void planeEnergyKernel(PlaneArgs<real> a) {
    const int lane = threadIdx.x & 63;
    const int wi = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (wi >= a.nWaves)
        return; /* no barriers below */
    const int yg = wi % a.nYG;
    const int t2 = wi / a.nYG;
    const int it = t2 % a.outTiles, kc = t2 / a.outTiles;
    const int kb0 = kc * a.chunk;
    const int kb1 = min(kb0 + a.chunk, a.kBlocks);
    const int g = lane >> 4, col = lane & 15;
    const int y0 = yg * TY * 16;
    const long long planeStride = (long long)a.outTiles * a.kBlocks * kWave;
    const uint4 *fr = a.frag + (long long)it * a.kBlocks * kWave + lane;

    i32x4 acc[P][TY];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int t = 0; t < TY; ++t)
            acc[p][t] = i32x4{0, 0, 0, 0};

    /* per y-tile: the lane's slice row of spin words (past m: row 0, unused) */
    const uint32_t *qrow[TY];
#pragma unroll
    for (int t = 0; t < TY; ++t) {
        const int y = y0 + 16 * t + col;
        qrow[t] = a.Qin + (long long)(y < a.m ? y : 0) * a.NwIn + (g >> 1);
    }
    const int qshift = 16 * (g & 1);

    for (int kb = kb0; kb < kb1; ++kb) {
        i32x4 A[P];
#pragma unroll
        for (int p = 0; p < P; ++p)
            A[p] = __builtin_bit_cast(i32x4, fr[p * planeStride + (long long)kb * kWave]);
        uint32_t bits[TY];
#pragma unroll
        for (int t = 0; t < TY; ++t)
            bits[t] = qrow[t][2 * kb];
#pragma unroll
        for (int t = 0; t < TY; ++t) {
            const i32x4 Bq = spinsI8((bits[t] >> qshift) & 0xFFFFu);
#pragma unroll
            for (int p = 0; p < P; ++p)
                acc[p][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[p], Bq, acc[p][t], 0, 0, 0);
        }
    }

    /* element (row 16 it + 4 g + r, slice y0 + 16 t + col) is acc[.][t][r] */
    const int ib = 16 * it + 4 * g;
#pragma unroll
    for (int t = 0; t < TY; ++t) {
        const int y = y0 + 16 * t + col;
        const uint32_t qb = (y < a.m) ? (a.Qout[(long long)y * a.NwOut + (ib >> 5)] >> (ib & 31)) & 15u : 0u;
        double s = 0.;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = ib + r;
            double F = (double)acc[P - 1][t][r];
#pragma unroll
            for (int p = P - 2; p >= 0; --p)
                F = F * 256. + (double)acc[p][t][r];
            F *= a.scale;
            if (i < a.nOut) {
                if (kc == 0)
                    F += (double)a.hOut[i];
                s += ((qb >> r) & 1u) ? F : -F;
            }
        }
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (g == 0 && y < a.m)
            a.partial[(long long)y * a.nSlots + (long long)kc * a.outTiles + it] = s;
    }
}

/* E[y] = -c - (sum of the slice's partials + hIn . q_in,y), fixed order */
template <class real>
__global__ __launch_bounds__(kFinalThreads) void planeFinalKernel(const double *partial, int nSlots, const real *hIn,
                                                                  const uint32_t *Qin, int NwIn, int K, double c,
                                                                  double *E) {
    __shared__ double red[kFinalThreads];
    const int y = blockIdx.x;
    double s = 0.;
    for (int j = threadIdx.x; j < nSlots; j += kFinalThreads)
        s += partial[(long long)y * nSlots + j];
    if (hIn) {
        for (int j = threadIdx.x; j < K; j += kFinalThreads) {
            const bool qp = (Qin[(long long)y * NwIn + (j >> 5)] >> (j & 31)) & 1u;
            s += qp ? (double)hIn[j] : -(double)hIn[j];
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kFinalThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        E[y] = -c - red[0];
}

/* y-tiles per wave: TY * P accumulator tiles (4 VGPRs each) stay within 192 */
int planeTilesY(int P) { return P <= 3 ? 16 : (P <= 6 ? 8 : 4); }

struct PlaneGrid {
    int TY, nYG, KS, chunk, nWaves, nSlots;
};

PlaneGrid planeGrid(const EnergyPlaneLayout &L, int m) {
    PlaneGrid gr;
    gr.TY = planeTilesY(L.P);
    gr.nYG = divru(m, 16 * gr.TY);
    const int waves0 = L.outTiles * gr.nYG;
    /* split the k-range until about 2048 waves run (8 XCDs x 32 CUs x 4
     * SIMDs, two waves each), keeping at least 4 k-blocks per chunk */
    int KS = divru(2048, waves0);
    KS = std::min(KS, std::max(1, L.kBlocks / 4));
    KS = std::max(KS, 1);
    gr.chunk = divru(L.kBlocks, KS);
    gr.KS = divru(L.kBlocks, gr.chunk);
    gr.nWaves = L.outTiles * gr.nYG * gr.KS;
    gr.nSlots = L.outTiles * gr.KS;
    return gr;
}

template <class real, int P, int TY>
void launchPlanes(const PlaneArgs<real> &a, hipStream_t stream) {
    planeEnergyKernel<real, P, TY><<<divru(a.nWaves, kWavesPerBlock), kWavesPerBlock * kWave, 0, stream>>>(a);
}

}  // namespace

template <class real>
EnergyPlaneLayout energyPlaneLayout(const real *d_B, long long sRow, long long sCol, int nOut, int K, int *d_scan,
                                    hipStream_t stream) {
    SQ_THROW_IF(nOut <= 0 || K <= 0, "energy planes: empty matrix");
    const int init[3] = {INT_MIN, INT_MAX, 0};
    HIP_CHECK(hipMemcpyAsync(d_scan, init, sizeof(init), hipMemcpyHostToDevice, stream));
    planeScanKernel<real><<<std::min(nOut, 2048), kScanThreads, 0, stream>>>(d_B, sRow, sCol, nOut, K, d_scan);
    HIP_CHECK(hipGetLastError());
    int got[3];
    HIP_CHECK(hipMemcpyAsync(got, d_scan, sizeof(got), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    EnergyPlaneLayout L;
    L.nOut = nOut;
    L.K = K;
    L.outTiles = divru(nOut, 16);
    L.kBlocks = divru(K, 64);
    L.finite = got[2] ? 0 : 1;
    const int pMax = sizeof(real) == 4 ? 6 : 7;
    if (got[0] == INT_MIN) { /* all zero */
        L.P = 1;
        L.es = 0;
        L.exact = 1;
        return L;
    }
    /* |V| < 2^(hi + 1 - es) <= 2^(8P - 2) keeps V within P balanced digits */
    const int bits = got[0] + 1 - got[1];
    const int P = (bits + 2 + 7) / 8;
    if (P <= pMax) {
        L.P = P;
        L.es = got[1];
        L.exact = 1;
    } else {
        L.P = pMax;
        L.es = got[0] + 3 - 8 * pMax;
        L.exact = 0;
    }
    return L;
}

size_t energyPlaneFragCount(const EnergyPlaneLayout &L) {
    return (size_t)L.P * L.outTiles * L.kBlocks * kWave;
}

size_t energyPlanePartialCount(const EnergyPlaneLayout &L, int m) {
    return (size_t)m * planeGrid(L, m).nSlots;
}

template <class real>
void energyPlaneBuild(const real *d_B, long long sRow, long long sCol, const EnergyPlaneLayout &L, void *d_frag,
                      hipStream_t stream) {
    SQ_THROW_IF(L.P < 1 || L.P > 7, "energy planes: bad plane count %d", L.P);
    const long long perPlane = (long long)L.outTiles * L.kBlocks * kWave;
    planeBuildKernel<real><<<(unsigned)((perPlane + 255) / 256), 256, 0, stream>>>(
        d_B, sRow, sCol, L.nOut, L.K, L.outTiles, L.kBlocks, L.P, L.es, (uint4 *)d_frag);
    HIP_CHECK(hipGetLastError());
}

template <class real>
void energyPlanes(const EnergyPlaneLayout &L, const void *d_frag, const real *d_hOut, const real *d_hIn, double c,
                  const uint32_t *d_Qin, int NwIn, const uint32_t *d_Qout, int NwOut, int m, double *d_partial,
                  double *d_E, hipStream_t stream) {
    SQ_THROW_IF(m <= 0, "energy planes: no slices");
    SQ_THROW_IF(2 * L.kBlocks > NwIn, "energy planes: spin rows shorter than the k-range");
    SQ_THROW_IF(divru(16 * L.outTiles, 32) > NwOut, "energy planes: spin rows shorter than the row range");
    const PlaneGrid gr = planeGrid(L, m);
    PlaneArgs<real> a;
    a.frag = (const uint4 *)d_frag;
    a.hOut = d_hOut;
    a.Qin = d_Qin;
    a.Qout = d_Qout;
    a.partial = d_partial;
    a.scale = std::ldexp(1., L.es);
    a.nOut = L.nOut;
    a.outTiles = L.outTiles;
    a.kBlocks = L.kBlocks;
    a.NwIn = NwIn;
    a.NwOut = NwOut;
    a.m = m;
    a.nYG = gr.nYG;
    a.KS = gr.KS;
    a.chunk = gr.chunk;
    a.nWaves = gr.nWaves;
    a.nSlots = gr.nSlots;
    switch (L.P) {
    case 1: launchPlanes<real, 1, 16>(a, stream); break;
    case 2: launchPlanes<real, 2, 16>(a, stream); break;
    case 3: launchPlanes<real, 3, 16>(a, stream); break;
    case 4: launchPlanes<real, 4, 8>(a, stream); break;
    case 5: launchPlanes<real, 5, 8>(a, stream); break;
    case 6: launchPlanes<real, 6, 8>(a, stream); break;
    case 7: launchPlanes<real, 7, 4>(a, stream); break;
    default: SQ_THROW("energy planes: bad plane count %d", L.P);
    }
    HIP_CHECK(hipGetLastError());
    planeFinalKernel<real><<<m, kFinalThreads, 0, stream>>>(d_partial, gr.nSlots, d_hIn, d_Qin, NwIn, L.K, c, d_E);
    HIP_CHECK(hipGetLastError());
}

template EnergyPlaneLayout energyPlaneLayout<float>(const float *, long long, long long, int, int, int *, hipStream_t);
template EnergyPlaneLayout energyPlaneLayout<double>(const double *, long long, long long, int, int, int *,
                                                     hipStream_t);
template void energyPlaneBuild<float>(const float *, long long, long long, const EnergyPlaneLayout &, void *,
                                      hipStream_t);
template void energyPlaneBuild<double>(const double *, long long, long long, const EnergyPlaneLayout &, void *,
                                       hipStream_t);
template void energyPlanes<float>(const EnergyPlaneLayout &, const void *, const float *, const float *, double,
                                  const uint32_t *, int, const uint32_t *, int, int, double *, double *, hipStream_t);
template void energyPlanes<double>(const EnergyPlaneLayout &, const void *, const double *, const double *, double,
                                   const uint32_t *, int, const uint32_t *, int, int, double *, double *,
                                   hipStream_t);

}  // namespace sqaod_rocm
