/*
 * annealOneStepNaive of the dense annealer (sqaodc/cpu/CPUDenseGraphAnnealer.cpp:
 * 268-280): N m attempts in one serial sequence, attempt t drawing its slice
 * y = randInt(m), then tryFlip's site x = randInt(N) and its uniform u
 * (:250-265).  Unlike the coloring step there is no structure to split
 * between slices: any attempt may touch any slice, so the step runs on one
 * workgroup, in speculative batches of 64 consecutive attempts.
 *
 *   - wave 0 evaluates attempts t .. t+63 (lane i: attempt t+i) against the
 *     state before the batch: local field F[y][x] (kept as in the coloring
 *     sweep, F[y][i] = sum_j J[i][j] q[y][j]), h[x], the spins of (y, x) and
 *     of the ring neighbours (y-1, x), (y+1, x), and the acceptance
 *     threshold in the reference's operation order (the expf of the CPU
 *     annealer, bit for bit, for float);
 *   - a ballot finds the first accepting lane a; lanes before a saw exactly
 *     the state the serial order gives them (no earlier attempt of the batch
 *     changed anything), lane a too, so attempts t .. t+a are decided;
 *   - the flip at (y_a, x_a) is applied by the whole workgroup (the spin bit
 *     and F[y_a][:] += -2 q J[x_a][:], the float operation of the coloring
 *     sweep's row update), and the next batch starts at t+a+1.
 *
 * Streams: with rng = mt19937 the words of the host generator in the
 * reference's order (y, x, u: three words per float attempt, four per
 * double); with rng = philox, attempt t of sweep s is the Philox4x32-10
 * block at counter (t_lo, t_hi, s_lo, s_hi ^ 0x20000000) under the seed key,
 * y = word 0 % m, x = word 1 % N, u from words 2 (and 3).
 *
 * Throughput is bounded by one workgroup's batch latency; the step is the
 * reference's alternative algorithm, kept for its semantics, while coloring
 * is the default and the fast path.
 */
#include "common.h"
#include "kernels.h"
#include "rng_device.h"
#include "sq_expf.h"

namespace sqaod_rocm {

namespace {

constexpr int kNaiveThreads = 1024;

template <class real>
struct NaiveDraw;
template <>
struct NaiveDraw<float> {
    static constexpr int words = 3;
    __device__ static float fromWords(const uint32_t *w) { return wordToUniform(w[2]); }
    __device__ static float fromBlock(const U4 &r) { return wordToUniform(r.z); }
    __device__ static float expo(float v, const uint64_t *tab) { return sq_expf_tab(v, tab); }
};
template <>
struct NaiveDraw<double> {
    static constexpr int words = 4;
    __device__ static double fromWords(const uint32_t *w) { return wordsToUniform(w[2], w[3]); }
    __device__ static double fromBlock(const U4 &r) { return wordsToUniform(r.z, r.w); }
    __device__ static double expo(double v, const uint64_t *) { return exp(v); }
};

__device__ __forceinline__ bool spinUp(const uint32_t *Q, int Nw, int y, int x) {
    return (Q[(long long)y * Nw + (x >> 5)] >> (x & 31)) & 1u;
}

template <class real, bool PHILOX>
__global__ __launch_bounds__(kNaiveThreads) void naiveSweepKernel(NaiveArgs<real> a) {
    __shared__ uint64_t expTab[32];
    __shared__ int sFirst, sY, sX;
    __shared__ real sD;
    if (threadIdx.x < 32)
        expTab[threadIdx.x] = sq_exp2_tab32[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const bool evaluator = threadIdx.x < 64;
    const long long total = (long long)a.N * a.m;
    unsigned long long flips = 0, batches = 0;
    long long t0 = 0;
    while (t0 < total) {
        if (evaluator) {
            const long long t = t0 + lane;
            const bool valid = t < total;
            int y = 0, x = 0;
            real u = (real)1.;
            if (valid) {
                if constexpr (PHILOX) {
                    const U4 r = philox4x32_10(U4{(uint32_t)t, (uint32_t)(t >> 32), a.s0, a.s1 ^ 0x20000000u},
                                               a.k0, a.k1);
                    y = (int)(r.x % (uint32_t)a.m);
                    x = (int)(r.y % (uint32_t)a.N);
                    u = NaiveDraw<real>::fromBlock(r);
                } else {
                    const uint32_t *w = a.words + t * NaiveDraw<real>::words;
                    y = (int)(w[0] % (uint32_t)a.m);
                    x = (int)(w[1] % (uint32_t)a.N);
                    u = NaiveDraw<real>::fromWords(w);
                }
            }
            const int n0 = (y == 0) ? a.m - 1 : y - 1;
            const int n1 = (y == a.m - 1) ? 0 : y + 1;
            const real f = a.F[(long long)y * a.Np + x];
            const real hx = a.h[x];
            const real q = spinUp(a.Qbits, a.Nw, y, x) ? (real)1. : (real)-1.;
            const real s0 = spinUp(a.Qbits, a.Nw, n0, x) ? (real)1. : (real)-1.;
            const real s1 = spinUp(a.Qbits, a.Nw, n1, x) ? (real)1. : (real)-1.;
            /* CPUDenseGraphAnnealer.cpp:258-262, the same operation order */
            real dE = a.twoDivM * q * (hx + (real)2. * f);
            dE -= q * (s0 + s1) * a.coef;
            const real e = NaiveDraw<real>::expo(-dE * a.beta, expTab);
            const real thr = (dE < (real)0.) ? (real)1. : e;
            const bool accept = valid && thr > u;
            const unsigned long long bal = __ballot(accept);
            const int first = bal ? __ffsll((long long)bal) - 1 : -1;
            if (lane == 0)
                sFirst = first;
            if (lane == first) {
                sY = y;
                sX = x;
                sD = (real)-2. * q;
            }
        }
        __syncthreads();
        const int first = sFirst;
        if (first >= 0) {
            const int y = sY, x = sX;
            const real d = sD;
            real *Fy = a.F + (long long)y * a.Np;
            const real *Jx = a.J + (long long)x * a.Np;
            for (int i = threadIdx.x; i < a.N; i += kNaiveThreads)
                Fy[i] = Fy[i] + d * Jx[i];
            if (threadIdx.x == 0)
                a.Qbits[(long long)y * a.Nw + (x >> 5)] ^= 1u << (x & 31);
            t0 += first + 1;
            ++flips;
        } else {
            t0 += 64;
        }
        ++batches;
        __syncthreads();
    }
    if (threadIdx.x == 0 && a.counters) {
        atomicAdd(&a.counters[0], flips);
        atomicAdd(&a.counters[1], batches);
    }
}

}  // namespace

template <class real>
void naiveSweep(const NaiveArgs<real> &a, hipStream_t stream) {
    SQ_THROW_IF(a.N <= 0 || a.m <= 0, "naive step: empty problem");
    if (a.words)
        naiveSweepKernel<real, false><<<1, kNaiveThreads, 0, stream>>>(a);
    else
        naiveSweepKernel<real, true><<<1, kNaiveThreads, 0, stream>>>(a);
    HIP_CHECK(hipGetLastError());
}

template void naiveSweep<float>(const NaiveArgs<float> &, hipStream_t);
template void naiveSweep<double>(const NaiveArgs<double> &, hipStream_t);

}  // namespace sqaod_rocm
