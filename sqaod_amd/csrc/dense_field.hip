/*
 * Spin-matrix product F = Q . B on the vector ALU (Q: m slices of spins as
 * bits, B: K x nCols, zero padding): the dense annealer's local-field
 * refresh (B = J, the GEMM the sweep's LDS fields are seeded from; the CPU
 * annealer recomputes the same J-row . q-row dot per attempt,
 * CPUDenseGraphAnnealer.cpp:257) and the bipartite half-steps' dEmat.
 *
 * This file holds the vector-ALU form (fp64 J for double solvers, and fp32 J
 * as the cross-check the MFMA kernels of dense_field_mfma.hip are tested
 * against) and the bf16 hi+lo split of J.  Each block owns 256 columns and a
 * tile of slices; spins enter as sign flips of J, so the products are exact.
 */
#include "common.h"
#include "kernels.h"
#include "bf16.h"

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

}  // namespace

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
