/*
 * Local-field refresh F = Q . J for the dense annealer (Q: m x N spins as
 * bits, J: N x Np symmetric, zero padding), the GEMM the sweep's LDS fields
 * are seeded from (the CPU annealer recomputes the same J-row . q-row dot per
 * attempt, CPUDenseGraphAnnealer.cpp:257).
 *
 * This file holds the vector-ALU forms (fp64 J for double solvers, and the
 * fp32 / bf16 hi+lo forms kept as the reference the MFMA kernels of
 * dense_field_mfma.hip are tested against).  Each block owns 256 columns and a tile of slices;
 * spins enter as sign flips of J, so the products are exact.
 */
#include "common.h"
#include "kernels.h"

namespace sqaod_rocm {

namespace {

constexpr int kCols = 256;
constexpr int kYTile = 32;

template <class real, int FMT>
__global__ __launch_bounds__(kCols) void fieldRefreshKernel(const real *J, const uint16_t *Jhi, const uint16_t *Jlo,
                                                            const uint32_t *Qbits, int N, int Np, int m, int Nw,
                                                            real *F) {
#pragma clang fp contract(off)
    __shared__ uint32_t qs[kYTile];
    const int i = blockIdx.x * kCols + threadIdx.x;
    const int y0 = blockIdx.y * kYTile;
    real acc[kYTile];
#pragma unroll
    for (int t = 0; t < kYTile; ++t)
        acc[t] = (real)0;
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
            const long long off = (long long)(j0 + jj) * Np + i;
            real v;
            if constexpr (FMT == fieldBF16x2)
                v = (real)(__uint_as_float((uint32_t)Jhi[off] << 16) + __uint_as_float((uint32_t)Jlo[off] << 16));
            else
                v = J[off];
#pragma unroll
            for (int t = 0; t < kYTile; ++t)
                acc[t] += ((qw[t] >> jj) & 1u) ? v : -v;
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < kYTile; ++t)
        if (y0 + t < m)
            F[(long long)(y0 + t) * Np + i] = acc[t];
}

__global__ void splitBF16Kernel(const float *J, size_t count, uint16_t *hi, uint16_t *lo, float *Jsplit) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count)
        return;
    float v = J[idx];
    /* round to nearest even into bf16 (finite inputs) */
    uint32_t b = __float_as_uint(v);
    uint32_t r = b + 0x7fffu + ((b >> 16) & 1u);
    uint16_t h = (uint16_t)(r >> 16);
    float hv = __uint_as_float((uint32_t)h << 16);
    float rem = v - hv; /* exact */
    uint32_t bl = __float_as_uint(rem);
    uint32_t rl = bl + 0x7fffu + ((bl >> 16) & 1u);
    hi[idx] = h;
    lo[idx] = (uint16_t)(rl >> 16);
    Jsplit[idx] = hv + __uint_as_float((rl >> 16) << 16);
}

}  // namespace

void splitBF16(const float *d_J, size_t count, uint16_t *d_hi, uint16_t *d_lo, float *d_Jsplit, hipStream_t stream) {
    if (!count)
        return;
    splitBF16Kernel<<<(unsigned)((count + 255) / 256), 256, 0, stream>>>(d_J, count, d_hi, d_lo, d_Jsplit);
    HIP_CHECK(hipGetLastError());
}

template <class real>
void fieldRefresh(const real *d_J, const uint32_t *d_Qbits, int N, int Np, int m, int Nw, real *d_F,
                  hipStream_t stream) {
    dim3 grid(Np / kCols, divru(m, kYTile));
    fieldRefreshKernel<real, fieldF32><<<grid, kCols, 0, stream>>>(d_J, nullptr, nullptr, d_Qbits, N, Np, m, Nw, d_F);
    HIP_CHECK(hipGetLastError());
}

void fieldRefreshBF16x2Vector(const uint16_t *d_hi, const uint16_t *d_lo, const uint32_t *d_Qbits, int N, int Np,
                              int m, int Nw, float *d_F, hipStream_t stream) {
    dim3 grid(Np / kCols, divru(m, kYTile));
    fieldRefreshKernel<float, fieldBF16x2><<<grid, kCols, 0, stream>>>(nullptr, d_hi, d_lo, d_Qbits, N, Np, m, Nw,
                                                                       d_F);
    HIP_CHECK(hipGetLastError());
}

template void fieldRefresh<float>(const float *, const uint32_t *, int, int, int, int, float *, hipStream_t);
template void fieldRefresh<double>(const double *, const uint32_t *, int, int, int, int, double *, hipStream_t);

}  // namespace sqaod_rocm
