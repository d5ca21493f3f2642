/*
 * FP64 energy recompute for the annealers, evaluated in double from the
 * stored real J and h, with a fixed reduction order so the result does not
 * depend on scheduling:
 *   dense     E_y = -c - sum_i q_yi (h_i + sum_j J_ij q_yj)
 *             (DGFuncs::calculate_E, sqaodc/cpu/SharedFormulas.cpp:69-83);
 *   bipartite E_y = -c - h0.q0_y - sum_i q1_yi (h1_i + sum_j J_ij q0_yj)
 *             (BGFuncs::calculate_E, SharedFormulas.cpp:184-202).
 * Both are sums over "outer" spins of (h + B^T . inner spins): the partial
 * kernel takes the two spin sets separately.  Also the Trotter coupling
 * sum_y q_y . q_(y+1)%m of getSystemE (CPUDenseGraphAnnealer.cpp:236-240).
 */
#include "common.h"
#include "kernels.h"

namespace sqaod_rocm {

namespace {

constexpr int kCols = 256;
constexpr int kYTile = 16;

/* partial[y][bx] = sum over outer spins i in block bx's 256 columns (i < nOut)
 * of qo_yi (h_i + sum_{j<K} B[j][i] qi_yj); B row-major with stride ldB. */
template <class real>
__global__ __launch_bounds__(kCols) void energyPartialKernel(const real *B, int K, int ldB, const real *h,
                                                             const uint32_t *Qin, int NwIn, const uint32_t *Qout,
                                                             int NwOut, int nOut, int m, double *partial) {
    /* block (bx, by): columns i in [bx*256, +256), slices [by*16, +16) */
    __shared__ uint32_t qs[kYTile];
    __shared__ double red[kYTile][kCols / 64];
    const int i = blockIdx.x * kCols + threadIdx.x;
    const int y0 = blockIdx.y * kYTile;
    double acc[kYTile];
#pragma unroll
    for (int t = 0; t < kYTile; ++t)
        acc[t] = 0.;
    for (int j0 = 0; j0 < K; j0 += 32) {
        if (threadIdx.x < kYTile)
            qs[threadIdx.x] = (y0 + (int)threadIdx.x < m) ? Qin[(long long)(y0 + threadIdx.x) * NwIn + (j0 >> 5)] : 0u;
        __syncthreads();
        uint32_t qw[kYTile];
#pragma unroll
        for (int t = 0; t < kYTile; ++t)
            qw[t] = qs[t];
        const int jn = (K - j0) < 32 ? (K - j0) : 32;
        for (int jj = 0; jj < jn; ++jj) {
            /* sum_j B_ji q_j read down column i (coalesced) */
            const double v = (double)B[(long long)(j0 + jj) * ldB + i];
#pragma unroll
            for (int t = 0; t < kYTile; ++t)
                acc[t] += ((qw[t] >> jj) & 1u) ? v : -v;
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool colOk = i < nOut;
    const double hi = colOk ? (double)h[i] : 0.;
#pragma unroll
    for (int t = 0; t < kYTile; ++t) {
        double v = 0.;
        if (colOk && y0 + t < m) {
            const bool qp = (Qout[(long long)(y0 + t) * NwOut + (i >> 5)] >> (i & 31)) & 1u;
            const double s = hi + acc[t];
            v = qp ? s : -s;
        }
        /* wave reduction in a fixed butterfly order */
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_xor(v, off);
        if (lane == 0)
            red[t][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < kYTile) {
        const int t = threadIdx.x;
        if (y0 + t < m) {
            double s = ((red[t][0] + red[t][1]) + (red[t][2] + red[t][3]));
            partial[(long long)(y0 + t) * gridDim.x + blockIdx.x] = s;
        }
    }
}

/* E[y] = -c - (sum of partial0[y][:]) - (sum of partial1[y][:]) */
__global__ void energyFinalKernel(const double *partial0, int nPart0, const double *partial1, int nPart1, int m,
                                  double c, double *E) {
    int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= m)
        return;
    double s = 0.;
    for (int b = 0; b < nPart0; ++b)
        s += partial0[(long long)y * nPart0 + b];
    double s1 = 0.;
    for (int b = 0; b < nPart1; ++b)
        s1 += partial1[(long long)y * nPart1 + b];
    E[y] = -c - s - s1;
}

__global__ void spinDotSumKernel(const uint32_t *Qbits, int N, int m, int Nw, double *out) {
    /* one block; q_y . q_y' = N - 2 popcount(bits_y xor bits_y') (padding bits are 0 in both) */
    __shared__ long long red[256];
    long long acc = 0;
    for (long long idx = threadIdx.x; idx < (long long)m * Nw; idx += blockDim.x) {
        int y = (int)(idx / Nw), w = (int)(idx % Nw);
        int y1 = (y + 1) % m;
        acc += __popc(Qbits[(long long)y * Nw + w] ^ Qbits[(long long)y1 * Nw + w]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[0] = (double)((long long)N * m - 2 * red[0]);
}

}  // namespace

template <class real>
void energyFP64(const real *d_J, const real *d_h, double c, const uint32_t *d_Qbits, int N, int Np, int m, int Nw,
                double *d_partial, double *d_E, hipStream_t stream) {
    dim3 grid(Np / kCols, divru(m, kYTile));
    energyPartialKernel<real><<<grid, kCols, 0, stream>>>(d_J, N, Np, d_h, d_Qbits, Nw, d_Qbits, Nw, N, m, d_partial);
    HIP_CHECK(hipGetLastError());
    energyFinalKernel<<<divru(m, 64), 64, 0, stream>>>(d_partial, (int)grid.x, nullptr, 0, m, c, d_E);
    HIP_CHECK(hipGetLastError());
}

template <class real>
void energyBipartiteFP64(const real *d_JT, const real *d_h0, const real *d_h1, double c, const uint32_t *d_Q0,
                         const uint32_t *d_Q1, int N0, int N1, int N0p, int N1p, int m, double *d_partial,
                         double *d_E, hipStream_t stream) {
    SQ_THROW_IF(N0p % kCols != 0 || N1p % kCols != 0, "bipartite energy: padded sizes must be multiples of %d", kCols);
    const int N0w = N0p / 32, N1w = N1p / 32;
    /* sum_i q1_i (h1_i + sum_j J_ij q0_j) = sum_i q1_i (h1_i + sum_j JT_ji q0_j) */
    dim3 g1(N1p / kCols, divru(m, kYTile));
    double *p1 = d_partial, *p0 = d_partial + (size_t)m * g1.x;
    energyPartialKernel<real><<<g1, kCols, 0, stream>>>(d_JT, N0, N1p, d_h1, d_Q0, N0w, d_Q1, N1w, N1, m, p1);
    HIP_CHECK(hipGetLastError());
    /* h0 . q0: the same kernel with an empty inner sum */
    dim3 g0(N0p / kCols, divru(m, kYTile));
    energyPartialKernel<real><<<g0, kCols, 0, stream>>>(d_JT, 0, N1p, d_h0, d_Q0, N0w, d_Q0, N0w, N0, m, p0);
    HIP_CHECK(hipGetLastError());
    energyFinalKernel<<<divru(m, 64), 64, 0, stream>>>(p1, (int)g1.x, p0, (int)g0.x, m, c, d_E);
    HIP_CHECK(hipGetLastError());
}

void spinDotSum(const uint32_t *d_Qbits, int N, int m, int Nw, double *d_out, hipStream_t stream) {
    spinDotSumKernel<<<1, 256, 0, stream>>>(d_Qbits, N, m, Nw, d_out);
    HIP_CHECK(hipGetLastError());
}

template void energyFP64<float>(const float *, const float *, double, const uint32_t *, int, int, int, int, double *,
                                double *, hipStream_t);
template void energyFP64<double>(const double *, const double *, double, const uint32_t *, int, int, int, int,
                                 double *, double *, hipStream_t);
template void energyBipartiteFP64<float>(const float *, const float *, const float *, double, const uint32_t *,
                                         const uint32_t *, int, int, int, int, int, double *, double *, hipStream_t);
template void energyBipartiteFP64<double>(const double *, const double *, const double *, double, const uint32_t *,
                                          const uint32_t *, int, int, int, int, int, double *, double *, hipStream_t);

}  // namespace sqaod_rocm

/* ---------------- system energy (dense getSystemE) ---------------- */

namespace sqaod_rocm {

namespace {

template <class real>
__global__ void systemEnergyKernel(const double *E, int m, real sgnE, const double *spinDot, int sqa, real coef,
                                   real sgnOut, real *out) {
    real s = (real)0.;
    for (int y = 0; y < m; ++y)
        s += sgnE * (real)E[y];
    s = s / (real)m;
    if (sqa)
        s -= (real)spinDot[0] * coef;
    out[0] = s * sgnOut;
}

}  // namespace

template <class real>
void systemEnergy(const double *d_E, int m, real sgnE, const double *d_spinDot, int sqa, real coef, real sgnOut,
                  real *d_out, hipStream_t stream) {
    systemEnergyKernel<real><<<1, 1, 0, stream>>>(d_E, m, sgnE, d_spinDot, sqa, coef, sgnOut, d_out);
    HIP_CHECK(hipGetLastError());
}

template void systemEnergy<float>(const double *, int, float, const double *, int, float, float, float *,
                                  hipStream_t);
template void systemEnergy<double>(const double *, int, double, const double *, int, double, double, double *,
                                   hipStream_t);

}  // namespace sqaod_rocm

/* ---------------- batched bilinear forms (the formulas objects) ---------------- */

namespace sqaod_rocm {

namespace {

constexpr int kQfThreads = 256;

/* One block per output: E[o] = cst + sgn (sum_j lin0[j] v[j] + sum_i u[i] (lin1[i] + sum_j A[i][j] v[j])),
 * u = U[iu[o]] (nu entries), v = V[iv[o]] (nv entries), int8 vectors; FP64
 * accumulation, fixed reduction order (per-thread strided sums, then a tree). */
template <class real>
__global__ __launch_bounds__(kQfThreads) void quadFormKernel(QuadFormArgs<real> a) {
    __shared__ double red[kQfThreads];
    const int o = blockIdx.x;
    const int bu = a.twoD ? o / a.nV : o;
    const int bv = a.twoD ? o % a.nV : o;
    const int8_t *u = a.U + (size_t)bu * a.nu;
    const int8_t *v = a.V + (size_t)bv * a.nv;
    double s = 0.;
    for (int i = threadIdx.x; i < a.nu; i += kQfThreads) {
        if (u[i] == 0)
            continue;
        double r = a.lin1 ? (double)a.lin1[i] : 0.;
        const real *Ai = a.A + (size_t)i * a.nv;
        for (int j = 0; j < a.nv; ++j)
            if (v[j])
                r += (double)Ai[j] * (double)v[j];
        s += (double)u[i] * r;
    }
    if (a.lin0)
        for (int j = threadIdx.x; j < a.nv; j += kQfThreads)
            s += (double)a.lin0[j] * (double)v[j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kQfThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        a.E[o] = (real)(a.cst + a.sgn * red[0]);
}

}  // namespace

template <class real>
void quadForm(const QuadFormArgs<real> &a, int nOut, hipStream_t stream) {
    if (nOut <= 0)
        return;
    SQ_THROW_IF(a.nu <= 0 || a.nv <= 0, "bilinear form: empty vectors");
    quadFormKernel<real><<<nOut, kQfThreads, 0, stream>>>(a);
    HIP_CHECK(hipGetLastError());
}

template void quadForm<float>(const QuadFormArgs<float> &, int, hipStream_t);
template void quadForm<double>(const QuadFormArgs<double> &, int, hipStream_t);

}  // namespace sqaod_rocm
