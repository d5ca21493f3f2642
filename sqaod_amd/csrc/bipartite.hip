/*
 * Flip phase of the bipartite annealer's half-steps
 * (CPUBipartiteGraphAnnealer.cpp:331-362 coloring, :485-530 SA coloring).
 *
 * Within a half-step the annealed part's spins do not interact with each
 * other: spin (im, iq) sees only the fixed part through dEmat[im][iq], its
 * field h[iq] and the same site of the two neighbouring Trotter slices.  The
 * CPU visits even slices, then odd slices, each site in order; since every
 * decision depends only on slices of the other colour, all sites of one
 * colour are decided at once.  A phase is one launch over (slice, 64-site
 * group): one wave per group, decisions collected with a ballot, each wave
 * rewriting the two spin words it owns.  With m odd, slices 0 and m-1 are
 * both even and adjacent, and the CPU reaches m-1 after 0, so m-1 is a phase
 * of its own between the even and the odd phase.
 *
 * The uniform of spin (im, iq) is the one the serial CPU order assigns it:
 * position rank(im) * N + iq of the half-step's stream (rank = position of
 * im in the visiting order), so an MT19937 replay matches sqaod.cpu bit for
 * bit; the Philox stream keys the same position.
 */
#include "common.h"
#include "kernels.h"
#include "rng_device.h"
#include "sq_expf.h"

namespace sqaod_rocm {

namespace {

constexpr int kWave = 64;

template <class real>
__device__ __forceinline__ real expNeg(real x);
template <>
__device__ __forceinline__ float expNeg<float>(float x) {
    return sq_expf(x);
}
template <>
__device__ __forceinline__ double expNeg<double>(double x) {
    return exp(x);
}

template <class real>
__global__ __launch_bounds__(kWave) void bgFlipKernel(BGFlipArgs<real> a) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x;
    const int iq = blockIdx.x * kWave + lane;
    const int im = a.slice0 + (int)blockIdx.y * a.sliceStep;
    const bool valid = iq < a.N;
    const long long wordIdx = (long long)im * a.Nw + (iq >> 5);
    const uint32_t myWord = a.Q[(long long)im * a.Nw + ((blockIdx.x * kWave + (lane & 32)) >> 5)];
    bool flip = false;
    if (valid) {
        const real q = ((myWord >> (iq & 31)) & 1u) ? real(1.) : real(-1.);
        const real dEv = a.dEmat[(long long)im * a.ldE + iq];
        real dE;
        if (a.sa) {
            /* tryFlipSA, :485-496 */
            dE = real(2.) * q * (a.h[iq] + dEv);
        } else {
            /* tryFlip, :331-344 (no factor 2 on dEmat) */
            const int n0 = (im + a.m - 1) % a.m, n1 = (im + 1) % a.m;
            const real q0 = ((a.Q[(long long)n0 * a.Nw + (iq >> 5)] >> (iq & 31)) & 1u) ? real(1.) : real(-1.);
            const real q1 = ((a.Q[(long long)n1 * a.Nw + (iq >> 5)] >> (iq & 31)) & 1u) ? real(1.) : real(-1.);
            dE = a.twoDivM * q * (a.h[iq] + dEv);
            dE -= q * (q0 + q1) * a.coef;
        }
        const real thresh = dE < real(0.) ? real(1.) : expNeg<real>(-dE * a.beta);
        /* position of this spin in the serial visiting order */
        const int rank = a.sa ? im : ((im & 1) ? a.nEven + (im >> 1) : (im >> 1));
        const long long pos = (long long)rank * a.N + iq;
        real u;
        if (a.words) {
            if (sizeof(real) == 4)
                u = (real)wordToUniform(a.words[pos]);
            else
                u = (real)wordsToUniform(a.words[2 * pos], a.words[2 * pos + 1]);
        } else {
            const U4 r = philox4x32_10(U4{(uint32_t)pos, (uint32_t)(pos >> 32), a.ctr0, a.ctr1}, a.key0, a.key1);
            u = (sizeof(real) == 4) ? (real)wordToUniform(r.x) : (real)wordsToUniform(r.x, r.y);
        }
        flip = thresh > u;
    }
    /* lanes 0..31 own word 2*bx, lanes 32..63 word 2*bx+1 of the row */
    const unsigned long long mask = __ballot(flip);
    if ((lane & 31) == 0 && (blockIdx.x * kWave + lane) < a.Nw * 32) {
        const uint32_t bits = (uint32_t)(mask >> (lane & 32));
        a.Q[wordIdx] = myWord ^ bits;
    }
}

}  // namespace

template <class real>
void bipartiteFlipPhase(const BGFlipArgs<real> &a, int nSlices, hipStream_t stream) {
    if (nSlices <= 0)
        return;
    SQ_THROW_IF(a.N <= 0 || a.Nw * 32 < a.N || a.Nw % 2 != 0, "bipartite flip: bad row of %d spins in %d words", a.N,
                a.Nw);
    SQ_THROW_IF(a.slice0 < 0 || a.slice0 + (nSlices - 1) * a.sliceStep >= a.m, "bipartite flip: slices out of range");
    dim3 grid((unsigned)divru(a.N, kWave), (unsigned)nSlices);
    bgFlipKernel<real><<<grid, kWave, 0, stream>>>(a);
    HIP_CHECK(hipGetLastError());
}

template void bipartiteFlipPhase<float>(const BGFlipArgs<float> &, int, hipStream_t);
template void bipartiteFlipPhase<double>(const BGFlipArgs<double> &, int, hipStream_t);

}  // namespace sqaod_rocm
