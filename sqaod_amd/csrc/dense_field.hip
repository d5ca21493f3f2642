/*
 * Spin-matrix product F = Q . B on the vector ALU (Q: m slices of spins as
 * bits, B: K x nCols, zero padding): the dense annealer's local-field
 * refresh (B = J, the GEMM the sweep's LDS fields are seeded from; the CPU
 * annealer recomputes the same J-row . q-row dot per attempt,
 * CPUDenseGraphAnnealer.cpp:257) and the bipartite half-steps' dEmat.
 *
 * This file holds the vector-ALU form (fp64 J for double solvers, and fp32 J
 * as the cross-check the MFMA kernels of dense_field_mfma.hip are tested
 * against), the bf16 hi+lo split of J and the FP8 (e4m3) quantization of J.
 * Each block owns 256 columns and a tile of slices; spins enter as sign flips
 * of J, so the products are exact.
 */
#include <climits>

#include "common.h"
#include "kernels.h"
#include "bf16.h"
#include "fp8.h"

namespace sqaod_rocm {

namespace {

constexpr int kCols = 256;
constexpr int kYTile = 32;

template <class real>
__global__ __launch_bounds__(kCols) void fieldRefreshKernel(const real *B, int K, int ldB, const uint32_t *Qbits, int m,
                                                            int Nw, int ldF, real *F) {
#pragma clang fp contract(off)
    __shared__ uint32_t qs[kYTile];
    const int i = blockIdx.x * kCols + threadIdx.x;
    const int y0 = blockIdx.y * kYTile;
    real acc[kYTile];
#pragma unroll
    for (int t = 0; t < kYTile; ++t)
        acc[t] = (real)0;
    for (int j0 = 0; j0 < K; j0 += 32) {
        if (threadIdx.x < kYTile)
            qs[threadIdx.x] = (y0 + (int)threadIdx.x < m) ? Qbits[(long long)(y0 + threadIdx.x) * Nw + (j0 >> 5)] : 0u;
        __syncthreads();
        uint32_t qw[kYTile];
#pragma unroll
        for (int t = 0; t < kYTile; ++t)
            qw[t] = qs[t];
        const int jn = (K - j0) < 32 ? (K - j0) : 32;
        for (int jj = 0; jj < jn; ++jj) {
            const real v = B[(long long)(j0 + jj) * ldB + i];
#pragma unroll
            for (int t = 0; t < kYTile; ++t)
                acc[t] += ((qw[t] >> jj) & 1u) ? v : -v;
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < kYTile; ++t)
        if (y0 + t < m)
            F[(long long)(y0 + t) * ldF + i] = acc[t];
}

__global__ void splitBF16Kernel(const float *J, size_t count, float *Jsplit) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count)
        return;
    uint16_t hi, lo;
    bf16Split(J[idx], hi, lo);
    Jsplit[idx] = bf16ToFloat(hi) + bf16ToFloat(lo); /* exact: lo holds the leading bits of J - hi */
}

/* FP8 rows, step 1: the exponent a_i of row i's scale, the smallest a with
 * max_j |J[i][j]| <= 448 * 2^a (448 = the largest e4m3 value); INT_MIN for
 * an all-zero row.  One block per row. */
__global__ __launch_bounds__(256) void fp8RowExpKernel(const float *J, int N, int ld, int *rowExp) {
    __shared__ float red[256];
    const int i = blockIdx.x;
    float mx = 0.f;
    for (int j = threadIdx.x; j < N; j += 256)
        mx = fmaxf(mx, fabsf(J[(long long)i * ld + j]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float r = red[0];
        int a = INT_MIN;
        if (r > 0.f) {
            int e;
            const float f = frexpf(r, &e); /* r = f 2^e, f in [0.5, 1): 448 2^(e-9) = 0.875 2^e */
            a = (f <= 0.875f) ? e - 9 : e - 8;
        }
        rowExp[i] = a;
    }
}

/* FP8 rows, step 2 (one block): row exponents are raised to within
 * kFP8ExpSpread of the largest, so every element of J' (below) is
 * representable in both its rows' grids; scale[i] = 2^a_i, 1 past N. */
constexpr int kFP8ExpSpread = 15;
__global__ __launch_bounds__(256) void fp8ClampExpKernel(int *rowExp, int N, int Np, float *scale) {
    __shared__ int red[256];
    int mx = INT_MIN;
    for (int i = threadIdx.x; i < N; i += 256)
        mx = max(mx, rowExp[i]);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    const int amax = red[0] == INT_MIN ? 0 : red[0];
    for (int i = threadIdx.x; i < Np; i += 256) {
        if (i < N) {
            const int a = max(rowExp[i], amax - kFP8ExpSpread);
            rowExp[i] = a;
            scale[i] = ldexpf(1.f, a);
        } else {
            scale[i] = 1.f;
        }
    }
}

/* FP8 rows, step 3: J'[i][j] = J[i][j] rounded to the e4m3 grid of the
 * coarser of its two rows' scales (2^max(a_i, a_j)), so J' is symmetric
 * whenever J is; byte[i][j] = J'[i][j] / 2^a_i exactly in e4m3.  Jq holds J'
 * (row stride ld), J8 the bytes (row stride ld8 >= ld), zero past N. */
__global__ void fp8QuantizeKernel(const float *J, int N, int ld, const int *rowExp, int ld8, uint8_t *J8,
                                  float *Jq) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)N * ld8)
        return;
    const int i = (int)(idx / ld8), j = (int)(idx % ld8);
    uint8_t byte = 0;
    float jp = 0.f;
    if (j < N) {
        const int ai = rowExp[i], aij = max(ai, rowExp[j]);
        const uint8_t q = floatToFp8(ldexpf(J[(long long)i * ld + j], -aij));
        jp = ldexpf(fp8ToFloat(q), aij);
        byte = floatToFp8(ldexpf(jp, -ai));
    }
    J8[idx] = byte;
    if (j < ld)
        Jq[(long long)i * ld + j] = jp;
}

}  // namespace

void quantizeFP8(const float *d_J, int N, int ld, int ld8, int *d_rowExp, float *d_scale, uint8_t *d_J8, float *d_Jq,
                 hipStream_t stream) {
    SQ_THROW_IF(N <= 0 || ld < N || ld8 < ld, "fp8 quantization: inconsistent shape");
    fp8RowExpKernel<<<N, 256, 0, stream>>>(d_J, N, ld, d_rowExp);
    HIP_CHECK(hipGetLastError());
    fp8ClampExpKernel<<<1, 256, 0, stream>>>(d_rowExp, N, ld, d_scale);
    HIP_CHECK(hipGetLastError());
    const long long n = (long long)N * ld8;
    fp8QuantizeKernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(d_J, N, ld, d_rowExp, ld8, d_J8, d_Jq);
    HIP_CHECK(hipGetLastError());
}

void splitBF16(const float *d_J, size_t count, float *d_Jsplit, hipStream_t stream) {
    if (!count)
        return;
    splitBF16Kernel<<<(unsigned)((count + 255) / 256), 256, 0, stream>>>(d_J, count, d_Jsplit);
    HIP_CHECK(hipGetLastError());
}

template <class real>
void fieldRefresh(const real *d_B, int K, int nCols, int ldB, const uint32_t *d_Qbits, int m, int Nw, int ldF,
                  real *d_F, hipStream_t stream) {
    SQ_THROW_IF(nCols <= 0 || nCols % kCols != 0, "field refresh: %d columns is not a multiple of %d", nCols, kCols);
    SQ_THROW_IF(ldB < nCols || ldF < nCols || K > Nw * 32, "field refresh: inconsistent strides");
    dim3 grid(nCols / kCols, divru(m, kYTile));
    fieldRefreshKernel<real><<<grid, kCols, 0, stream>>>(d_B, K, ldB, d_Qbits, m, Nw, ldF, d_F);
    HIP_CHECK(hipGetLastError());
}

template void fieldRefresh<float>(const float *, int, int, int, const uint32_t *, int, int, int, float *,
                                  hipStream_t);
template void fieldRefresh<double>(const double *, int, int, int, const uint32_t *, int, int, int, double *,
                                   hipStream_t);

}  // namespace sqaod_rocm
