/*
 * Device-side self checks exported for the test suite: an order-independent
 * checksum of sq_expf over a range of float bit patterns, so the device's
 * exp can be compared with the host C library's expf on every input without
 * copying 2^32 results back.
 */
#include "common.h"
#include "sq_expf.h"

namespace sqaod_rocm {

namespace {

__host__ __device__ inline unsigned long long mix64(unsigned long long z) {
    /* splitmix64 finaliser */
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void expfChecksumKernel(unsigned long long lo, unsigned long long hi, unsigned long long *out) {
    unsigned long long acc = 0;
    for (unsigned long long b = lo + (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; b < hi;
         b += (unsigned long long)gridDim.x * blockDim.x) {
        float r = sq_expf(__uint_as_float((uint32_t)b));
        uint32_t rb = __float_as_uint(r);
        if (r != r)
            rb = 0x7fc00000u;
        acc += mix64((b << 32) | rb);
    }
    for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0)
        atomicAdd(out, acc);
}

}  // namespace

unsigned long long expfChecksum(unsigned long long lo, unsigned long long hi) {
    unsigned long long *d = nullptr, h = 0;
    HIP_CHECK(hipMalloc((void **)&d, sizeof(unsigned long long)));
    HIP_CHECK(hipMemset(d, 0, sizeof(unsigned long long)));
    expfChecksumKernel<<<4096, 256>>>(lo, hi, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    HIP_CHECK(e);
    return h;
}

}  // namespace sqaod_rocm
