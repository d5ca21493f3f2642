/*
 * FP64 energy recompute for the dense annealer:
 *   E_y = -c - sum_i q_yi (h_i + sum_j J_ij q_yj)
 * (DGFuncs::calculate_E, sqaodc/cpu/SharedFormulas.cpp:69-83), evaluated in
 * double from the stored real J and h, with a fixed reduction order so the
 * result does not depend on scheduling.  Also the Trotter coupling
 * sum_y q_y . q_(y+1)%m of getSystemE (CPUDenseGraphAnnealer.cpp:236-240).
 */
#include "common.h"
#include "kernels.h"

namespace sqaod_rocm {

namespace {

constexpr int kCols = 256;
constexpr int kYTile = 16;

template <class real>
__global__ __launch_bounds__(kCols) void energyPartialKernel(const real *J, const real *h, const uint32_t *Qbits,
                                                             int N, int Np, int m, int Nw, double *partial) {
    /* block (bx, by): columns i in [bx*256, +256), slices [by*16, +16) */
    __shared__ uint32_t qs[kYTile];
    __shared__ double red[kYTile][kCols / 64];
    const int i = blockIdx.x * kCols + threadIdx.x;
    const int y0 = blockIdx.y * kYTile;
    double acc[kYTile];
#pragma unroll
    for (int t = 0; t < kYTile; ++t)
        acc[t] = 0.;
    for (int j0 = 0; j0 < N; j0 += 32) {
        if (threadIdx.x < kYTile)
            qs[threadIdx.x] = (y0 + (int)threadIdx.x < m) ? Qbits[(long long)(y0 + threadIdx.x) * Nw + (j0 >> 5)] : 0u;
        __syncthreads();
        uint32_t qw[kYTile];
#pragma unroll
        for (int t = 0; t < kYTile; ++t)
            qw[t] = qs[t];
        const int jn = (N - j0) < 32 ? (N - j0) : 32;
        for (int jj = 0; jj < jn; ++jj) {
            /* J symmetric: sum_j J_ij q_j read down column i (coalesced) */
            const double v = (double)J[(long long)(j0 + jj) * Np + i];
#pragma unroll
            for (int t = 0; t < kYTile; ++t)
                acc[t] += ((qw[t] >> jj) & 1u) ? v : -v;
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool colOk = i < N;
    const double hi = colOk ? (double)h[i] : 0.;
#pragma unroll
    for (int t = 0; t < kYTile; ++t) {
        double v = 0.;
        if (colOk && y0 + t < m) {
            const bool qp = (Qbits[(long long)(y0 + t) * Nw + (i >> 5)] >> (i & 31)) & 1u;
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

__global__ void energyFinalKernel(const double *partial, int nPart, int m, double c, double *E) {
    int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= m)
        return;
    double s = 0.;
    for (int b = 0; b < nPart; ++b)
        s += partial[(long long)y * nPart + b];
    E[y] = -c - s;
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
    energyPartialKernel<real><<<grid, kCols, 0, stream>>>(d_J, d_h, d_Qbits, N, Np, m, Nw, d_partial);
    HIP_CHECK(hipGetLastError());
    energyFinalKernel<<<divru(m, 64), 64, 0, stream>>>(d_partial, (int)grid.x, m, c, d_E);
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

}  // namespace sqaod_rocm
