/*
 * Random streams for the annealers.
 *
 *  - MT19937 on the device, continuing exactly the serial host generator of
 *    sqaodc/common/Random.cpp:64-149 (init_genrand, twist, tempering), so a
 *    run seeded like sqaod.cpu consumes the same words in the same order.
 *    One workgroup regenerates the 624-word block in four dependency phases
 *    (k < 227; 227 <= k < 454; 454 <= k < 623; k = 623).
 *  - Philox4x32-10 (Salmon et al., SC'11) as the counter-based stream of the
 *    fast path: attempt (slice y, iteration k) of sweep s is block
 *    (k, y, s_lo, s_hi) under key (seed_lo, seed_hi).
 *
 * Word -> (site, uniform) conversions follow Random.cpp:151-168:
 * site = word % N (modulo bias kept), float u = float(word) * 2^-32 (round to
 * nearest even, may equal 1.0f), double u = ((w1>>5)*2^26 + (w2>>6)) * 2^-53.
 */
#include "common.h"
#include "kernels.h"
#include "rng_device.h"

namespace sqaod_rocm {

namespace {

constexpr uint32_t kMatrixA = 0x9908b0dfu;
constexpr uint32_t kUpper = 0x80000000u;
constexpr uint32_t kLower = 0x7fffffffu;

__device__ __forceinline__ uint32_t mtMix(uint32_t a, uint32_t b) {
    uint32_t y = (a & kUpper) | (b & kLower);
    return (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
}

__device__ __forceinline__ uint32_t mtTemper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* one workgroup of 1024 threads continues st by n words into out */
__device__ void mtGenerateBlock(MTState *st, uint32_t *out, unsigned long long n) {
    __shared__ uint32_t mt[624];
    __shared__ uint32_t nw[624];
    const int t = threadIdx.x;
    if (t < 624)
        mt[t] = st->mt[t];
    uint32_t idx = st->idx;
    __syncthreads();
    unsigned long long pos = 0;
    while (pos < n) {
        if (idx >= 624) {
            /* sequential twist restated as four parallel phases */
            if (t < 227)
                nw[t] = mt[t + 397] ^ mtMix(mt[t], mt[t + 1]);
            __syncthreads();
            if (t >= 227 && t < 454)
                nw[t] = nw[t - 227] ^ mtMix(mt[t], mt[t + 1]);
            __syncthreads();
            if (t >= 454 && t < 623)
                nw[t] = nw[t - 227] ^ mtMix(mt[t], mt[t + 1]);
            __syncthreads();
            if (t == 623)
                nw[623] = nw[396] ^ mtMix(mt[623], nw[0]);
            __syncthreads();
            if (t < 624)
                mt[t] = nw[t];
            __syncthreads();
            idx = 0;
        }
        unsigned long long rem = n - pos;
        uint32_t cnt = (uint32_t)((rem < (unsigned long long)(624 - idx)) ? rem : (unsigned long long)(624 - idx));
        if ((uint32_t)t < cnt)
            out[pos + t] = mtTemper(mt[idx + t]);
        idx += cnt;
        pos += cnt;
    }
    __syncthreads();
    if (t < 624)
        st->mt[t] = mt[t];
    if (t == 0)
        st->idx = idx;
}

__global__ __launch_bounds__(1024) void mtGenerateKernel(MTState *st, uint32_t *out, unsigned long long n) {
    mtGenerateBlock(st, out, n);
}

/* Slices of one colour handled by worker w of T (libgomp's schedule(static)
 * without a chunk size: contiguous chunks, the first nHalf % T workers one
 * larger); *b = its first slice index within the colour. */
__device__ __forceinline__ int workerChunk(int nHalf, int T, int w, int *b) {
    const int q = nHalf / T, r = nHalf % T;
    *b = w * q + (w < r ? w : r);
    return q + (w < r ? 1 : 0);
}

/* Attempts worker w makes per iteration of the coloring step: its even
 * slices, its odd slices, and for odd m worker 0 also slice m-1
 * (CPUDenseGraphAnnealer.cpp:303-327). */
__device__ __forceinline__ int workerAttemptsPerIt(int nHalf, int T, int oddM, int w) {
    int b;
    return 2 * workerChunk(nHalf, T, w, &b) + ((w == 0 && oddM) ? 1 : 0);
}

/* block w: worker w's words of one coloring step, from its own generator */
__global__ __launch_bounds__(1024) void mtGenerateWorkersKernel(MTState *st, uint32_t *out, unsigned long long stride,
                                                                int T, int nHalf, int oddM, int N, int perAttempt) {
    const int w = blockIdx.x;
    const unsigned long long n =
        (unsigned long long)N * (unsigned long long)workerAttemptsPerIt(nHalf, T, oddM, w) * perAttempt;
    mtGenerateBlock(st + w, out + (unsigned long long)w * stride, n);
}

template <class real>
struct Draws;
template <>
struct Draws<float> {
    static constexpr int perAttempt = 2;
    __device__ static float uniform(const uint32_t *w) { return wordToUniform(w[1]); }
    __device__ static float uniform(const U4 &r) { return wordToUniform(r.y); }
};
template <>
struct Draws<double> {
    static constexpr int perAttempt = 3;
    __device__ static double uniform(const uint32_t *w) { return wordsToUniform(w[1], w[2]); }
    __device__ static double uniform(const U4 &r) { return wordsToUniform(r.y, r.z); }
};

__device__ __forceinline__ long long attemptPosition(int order, int N, int m, int y, int k) {
    if (order == orderSliceMajor)
        return (long long)y * N + k;
    int nEven = (m + 1) / 2;
    int pos = (y & 1) ? nEven + (y >> 1) : (y >> 1);
    return (long long)k * m + pos;
}

/* Attempt (y, k) of a coloring step in sqaod.cpu's OpenMP order: the
 * worker owning slice y draws it at position k * (its attempts per
 * iteration) + (its position within the iteration: even chunk, then slice
 * m-1 for worker 0 and odd m, then odd chunk). */
template <class real>
__global__ void workerStreamsToAttemptsKernel(const uint32_t *words, unsigned long long stride, int N, int m, int T,
                                              uint32_t *X, real *U) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)N * m)
        return;
    const int y = (int)(idx / N), k = (int)(idx % N);
    const int nHalf = m / 2, oddM = m & 1;
    int w, j, b;
    if (oddM && y == m - 1) {
        w = 0;
        j = workerChunk(nHalf, T, 0, &b);
    } else {
        const int kk = y >> 1, q = nHalf / T, r = nHalf % T;
        w = (kk < r * (q + 1)) ? kk / (q + 1) : r + (kk - r * (q + 1)) / q;
        const int c = workerChunk(nHalf, T, w, &b);
        j = (kk - b) + ((y & 1) ? c + ((w == 0 && oddM) ? 1 : 0) : 0);
    }
    const long long p = (long long)k * workerAttemptsPerIt(nHalf, T, oddM, w) + j;
    const uint32_t *wd = words + (unsigned long long)w * stride + (unsigned long long)p * Draws<real>::perAttempt;
    X[idx] = wd[0] % (uint32_t)N;
    U[idx] = Draws<real>::uniform(wd);
}

template <class real>
__global__ void streamToAttemptsKernel(const uint32_t *words, int N, int m, int order, uint32_t *X, real *U) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)N * m;
    if (idx >= total)
        return;
    int y = (int)(idx / N), k = (int)(idx % N);
    long long a = attemptPosition(order, N, m, y, k);
    const uint32_t *w = words + a * Draws<real>::perAttempt;
    X[idx] = w[0] % (uint32_t)N;
    U[idx] = Draws<real>::uniform(w);
}

/* attempt idx = y N + k of a sweep: site and uniform from Philox keyed
 * (seed) at counter (k, y, sweep index) */
template <class real>
__device__ __forceinline__ void philoxAttempt(uint32_t k0, uint32_t k1, uint32_t s0, uint32_t s1, int N, long long idx,
                                              uint32_t *X, real *U) {
    int y = (int)(idx / N), k = (int)(idx % N);
    U4 r = philox4x32_10(U4{(uint32_t)k, (uint32_t)y, s0, s1}, k0, k1);
    X[idx] = r.x % (uint32_t)N;
    U[idx] = Draws<real>::uniform(r);
}

template <class real>
__global__ void philoxAttemptsKernel(uint32_t k0, uint32_t k1, uint32_t s0, uint32_t s1, int N, int m, uint32_t *X,
                                     real *U) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)N * m)
        return;
    philoxAttempt<real>(k0, k1, s0, s1, N, idx, X, U);
}

/* many sweeps' streams in one launch: job j owns blocks [blockStart[j], ...) */
template <class real>
__global__ void philoxAttemptsGroupedKernel(const AttemptStreamJob *jobs, const int *blockStart, int nJobs) {
    int lo = 0, hi = nJobs - 1;
    const int b = blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blockStart[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const AttemptStreamJob j = jobs[lo];
    const long long idx = (long long)(b - blockStart[lo]) * blockDim.x + threadIdx.x;
    if (idx >= (long long)j.N * j.m)
        return;
    philoxAttempt<real>((uint32_t)j.seed, (uint32_t)(j.seed >> 32), (uint32_t)j.sweepIdx, (uint32_t)(j.sweepIdx >> 32),
                        j.N, idx, j.X, reinterpret_cast<real *>(j.U));
}

__global__ void spinsFromWordsKernel(const uint32_t *words, long long rowStride, int N, int m, int Nw, uint32_t *Q) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (w >= Nw || y >= m)
        return;
    uint32_t bits = 0;
    for (int b = 0; b < 32; ++b) {
        int x = w * 32 + b;
        if (x < N && (words[(long long)y * rowStride + x] & 1u))
            bits |= 1u << b;
    }
    Q[(long long)y * Nw + w] = bits;
}

__global__ void spinsFromPhiloxKernel(uint32_t k0, uint32_t k1, uint32_t salt, int N, int m, int Nw, uint32_t *Q) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (w >= Nw || y >= m)
        return;
    uint32_t bits = 0;
    for (int g = 0; g < 8; ++g) {
        U4 r = philox4x32_10(U4{(uint32_t)(w * 8 + g), (uint32_t)y, salt, 0xffffffffu}, k0, k1);
        uint32_t v[4] = {r.x, r.y, r.z, r.w};
        for (int b = 0; b < 4; ++b) {
            int x = w * 32 + g * 4 + b;
            if (x < N && (v[b] & 1u))
                bits |= 1u << (g * 4 + b);
        }
    }
    Q[(long long)y * Nw + w] = bits;
}

__global__ void packSpinsKernel(const int8_t *q, int N, int m, int Nw, uint32_t *Q) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (w >= Nw || y >= m)
        return;
    uint32_t bits = 0;
    for (int b = 0; b < 32; ++b) {
        int x = w * 32 + b;
        if (x < N && q[(long long)y * N + x] > 0)
            bits |= 1u << b;
    }
    Q[(long long)y * Nw + w] = bits;
}

__global__ void unpackSpinsKernel(const uint32_t *Q, int N, int m, int Nw, int8_t *q) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)N * m)
        return;
    int y = (int)(idx / N), x = (int)(idx % N);
    uint32_t bits = Q[(long long)y * Nw + (x >> 5)];
    q[idx] = ((bits >> (x & 31)) & 1u) ? (int8_t)1 : (int8_t)-1;
}

inline unsigned gridFor(long long total, int block) { return (unsigned)((total + block - 1) / block); }

}  // namespace

void mtSeedHost(MTState *st, unsigned long long seed) {
    st->mt[0] = (uint32_t)(seed & 0xffffffffull);
    for (int i = 1; i < 624; ++i)
        st->mt[i] = 1812433253u * (st->mt[i - 1] ^ (st->mt[i - 1] >> 30)) + (uint32_t)i;
    st->idx = 624;
}

void mtGenerate(MTState *d_state, uint32_t *d_out, size_t n, hipStream_t stream) {
    if (n == 0)
        return;
    mtGenerateKernel<<<1, 1024, 0, stream>>>(d_state, d_out, (unsigned long long)n);
    HIP_CHECK(hipGetLastError());
}

unsigned long long mtWorkerStride(int N, int m, int T, int perAttempt) {
    const int nHalf = m / 2;
    const int cMax = nHalf / T + (nHalf % T ? 1 : 0);
    return (unsigned long long)N * (unsigned long long)(2 * cMax + 1) * (unsigned long long)perAttempt;
}

template <class real>
void workerStreamsToAttempts(MTState *d_states, uint32_t *d_words, int N, int m, int T, uint32_t *d_X, real *d_U,
                             hipStream_t stream) {
    SQ_THROW_IF(T < 1 || T > 1024, "worker count %d out of range", T);
    const int per = sizeof(real) == 4 ? 2 : 3;
    const unsigned long long stride = mtWorkerStride(N, m, T, per);
    mtGenerateWorkersKernel<<<T, 1024, 0, stream>>>(d_states, d_words, stride, T, m / 2, m & 1, N, per);
    HIP_CHECK(hipGetLastError());
    long long total = (long long)N * m;
    workerStreamsToAttemptsKernel<real><<<gridFor(total, 256), 256, 0, stream>>>(d_words, stride, N, m, T, d_X, d_U);
    HIP_CHECK(hipGetLastError());
}
template void workerStreamsToAttempts<float>(MTState *, uint32_t *, int, int, int, uint32_t *, float *, hipStream_t);
template void workerStreamsToAttempts<double>(MTState *, uint32_t *, int, int, int, uint32_t *, double *,
                                              hipStream_t);

template <class real>
void streamToAttempts(const uint32_t *d_words, int N, int m, int order, uint32_t *d_X, real *d_U, hipStream_t stream) {
    long long total = (long long)N * m;
    streamToAttemptsKernel<real><<<gridFor(total, 256), 256, 0, stream>>>(d_words, N, m, order, d_X, d_U);
    HIP_CHECK(hipGetLastError());
}

template <class real>
void philoxAttemptsGrouped(const AttemptStreamJob *d_jobs, const int *d_blockStart, int nJobs, int totalBlocks,
                           hipStream_t stream) {
    SQ_THROW_IF(nJobs <= 0 || totalBlocks <= 0, "grouped attempt streams: no work");
    philoxAttemptsGroupedKernel<real><<<totalBlocks, kAttemptBlock, 0, stream>>>(d_jobs, d_blockStart, nJobs);
    HIP_CHECK(hipGetLastError());
}
template void philoxAttemptsGrouped<float>(const AttemptStreamJob *, const int *, int, int, hipStream_t);
template void philoxAttemptsGrouped<double>(const AttemptStreamJob *, const int *, int, int, hipStream_t);

template <class real>
void philoxAttempts(unsigned long long seed, unsigned long long sweepIdx, int N, int m, uint32_t *d_X, real *d_U,
                    hipStream_t stream) {
    long long total = (long long)N * m;
    philoxAttemptsKernel<real><<<gridFor(total, 256), 256, 0, stream>>>(
        (uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)sweepIdx, (uint32_t)(sweepIdx >> 32), N, m, d_X, d_U);
    HIP_CHECK(hipGetLastError());
}

void spinsFromWords(const uint32_t *d_words, long long rowStride, int N, int m, int Nw, uint32_t *d_Qbits,
                    hipStream_t stream) {
    SQ_THROW_IF(rowStride < N, "spinsFromWords: row stride %lld < N=%d", rowStride, N);
    dim3 grid(gridFor(Nw, 64), m);
    spinsFromWordsKernel<<<grid, 64, 0, stream>>>(d_words, rowStride, N, m, Nw, d_Qbits);
    HIP_CHECK(hipGetLastError());
}

void spinsFromPhilox(unsigned long long seed, unsigned long long salt, int N, int m, int Nw, uint32_t *d_Qbits,
                     hipStream_t stream) {
    dim3 grid(gridFor(Nw, 64), m);
    spinsFromPhiloxKernel<<<grid, 64, 0, stream>>>((uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)salt, N, m, Nw,
                                                   d_Qbits);
    HIP_CHECK(hipGetLastError());
}

void packSpins(const int8_t *d_q, int N, int m, int Nw, uint32_t *d_Qbits, hipStream_t stream) {
    dim3 grid(gridFor(Nw, 64), m);
    packSpinsKernel<<<grid, 64, 0, stream>>>(d_q, N, m, Nw, d_Qbits);
    HIP_CHECK(hipGetLastError());
}

void unpackSpins(const uint32_t *d_Qbits, int N, int m, int Nw, int8_t *d_q, hipStream_t stream) {
    long long total = (long long)N * m;
    unpackSpinsKernel<<<gridFor(total, 256), 256, 0, stream>>>(d_Qbits, N, m, Nw, d_q);
    HIP_CHECK(hipGetLastError());
}

template void streamToAttempts<float>(const uint32_t *, int, int, int, uint32_t *, float *, hipStream_t);
template void streamToAttempts<double>(const uint32_t *, int, int, int, uint32_t *, double *, hipStream_t);
template void philoxAttempts<float>(unsigned long long, unsigned long long, int, int, uint32_t *, float *,
                                    hipStream_t);
template void philoxAttempts<double>(unsigned long long, unsigned long long, int, int, uint32_t *, double *,
                                     hipStream_t);

}  // namespace sqaod_rocm
