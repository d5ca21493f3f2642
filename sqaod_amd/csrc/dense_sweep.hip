/*
 * Dense-graph SQA sweep: one annealOneStep (N attempts per Trotter slice)
 * with the exact semantics of the serial coloring annealer
 * (sqaodc/cpu/CPUDenseGraphAnnealer.cpp:250-301: for each iteration, even
 * slices then odd slices, one Metropolis attempt per slice), or of the SA
 * annealer (:341-371) when sa != 0.
 *
 * Layout.  One wavefront owns one slice y for the whole sweep.  Its local
 * fields F[y][:] (= J . q_y, fp32/fp64), its spins (bits) and private
 * mirrors of the two neighbouring slices' spins live in LDS; h is shared by
 * the workgroup.  Attempt (y, k) draws site X[y][k] and uniform U[y][k]
 * (pre-generated streams).
 *
 * Speculative batches.  The sites and uniforms of the next 64 attempts are
 * known, so lane i evaluates attempt k+i assuming attempts k..k+i-1 flip
 * nothing.  A ballot finds the first accepting lane a: attempts k..k+a-1 are
 * confirmed rejections, attempt k+a is a confirmed flip, and the batch is cut
 * after it.  Acceptance rates of the annealing schedules are a few percent, so
 * most batches commit tens of attempts.  A flip updates F[y][:] with one J row.
 *
 * Trotter coupling without grid barriers.  Each slice publishes its
 * decisions D[y][k] (tag << 1 | flipped) with device-scope stores.  A slice
 * needs neighbour n's spin at site x as of the K attempts of n that precede
 * (y, k) in the serial order (K = k or k + 1).  It replays n's decided
 * prefix into its mirror; undecided neighbour attempts that precede a lane's
 * attempt and touch the same site block that lane (and the batch is cut
 * there).  Every wait is on a serially earlier attempt, so the waits form no
 * cycle and all waves finish; a spin limit turns a bug into an error instead
 * of a hang.  All waves must be co-resident: the launch is cooperative.
 */
#include "common.h"
#include "kernels.h"
#include "sq_expf.h"

namespace sqaod_rocm {

namespace {

constexpr int kWave = 64;
constexpr int kTblMax = 4096; /* entries of the pending-site table per slice */

template <class real>
struct Ops;
template <>
struct Ops<float> {
    /* the CPU annealer's expf(-dE*beta), bit-identical (sq_expf.h) */
    __device__ static float expo(float v) { return sq_expf(v); }
};
template <>
struct Ops<double> {
    __device__ static double expo(double v) { return exp(v); }
};

__device__ __forceinline__ uint32_t loadShared32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void storeShared32(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int firstLane(unsigned long long mask) {
    return mask ? (int)__builtin_ctzll(mask) : kWave;
}

/* entries of the per-slice pending-site table: a power of two >= min(Np, 4096);
 * sites are hashed by their low bits, a collision only delays a lane. */
__host__ __device__ inline int pendingTableSize(int Np) {
    int lim = Np < kTblMax ? Np : kTblMax;
    int t = 256;
    while (t < lim)
        t <<= 1;
    return t;
}

__device__ __forceinline__ int slicePos(int y, int m) {
    int nEven = (m + 1) / 2;
    return (y & 1) ? nEven + (y >> 1) : (y >> 1);
}

/* Row of J in the sweep's field format, 4 consecutive elements. */
template <class real, int FMT>
struct JRow;
template <class real>
struct JRow<real, fieldF32> {
    __device__ static void load4(const SweepArgs<real> &a, long long off, real v[4]) {
        const real *p = a.J + off;
        if constexpr (sizeof(real) == 4) {
            float4 t = *reinterpret_cast<const float4 *>(p);
            v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
        } else {
            double2 t0 = *reinterpret_cast<const double2 *>(p);
            double2 t1 = *reinterpret_cast<const double2 *>(p + 2);
            v[0] = t0.x; v[1] = t0.y; v[2] = t1.x; v[3] = t1.y;
        }
    }
};
template <class real>
struct JRow<real, fieldBF16x2> {
    __device__ static void load4(const SweepArgs<real> &a, long long off, real v[4]) {
        uint2 hi = *reinterpret_cast<const uint2 *>(a.Jhi + off);
        uint2 lo = *reinterpret_cast<const uint2 *>(a.Jlo + off);
        uint32_t h[4] = {hi.x << 16, hi.x & 0xffff0000u, hi.y << 16, hi.y & 0xffff0000u};
        uint32_t l[4] = {lo.x << 16, lo.x & 0xffff0000u, lo.y << 16, lo.y & 0xffff0000u};
#pragma unroll
        for (int t = 0; t < 4; ++t)
            v[t] = (real)(__uint_as_float(h[t]) + __uint_as_float(l[t])); /* exact: 16 significant bits */
    }
};

template <class real, bool PLAIN, int FMT>
__global__ void denseSweepKernel(SweepArgs<real> a, int slicesPerBlock, int xcdRemap) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char lds[];
    const int N = a.N, Np = a.Np, m = a.m, Nw = a.Nw;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int tblN = pendingTableSize(Np);
    const int tblMask = tblN - 1;

    /* XCD-aware placement: consecutive slices on the same XCD (blocks are
     * dispatched round-robin over the 8 XCDs). */
    int b = blockIdx.x;
    if (xcdRemap) {
        int per = gridDim.x / 8;
        b = (blockIdx.x % 8) * per + blockIdx.x / 8;
    }
    const int y = b * slicesPerBlock + w;

    /* LDS carve-up */
    real *hS = reinterpret_cast<real *>(lds);
    unsigned char *p = lds + (size_t)Np * sizeof(real);
    const size_t perWave = (PLAIN ? 0 : (size_t)Np * sizeof(real)) + (size_t)tblN * 4 + (size_t)3 * Nw * 4;
    p += perWave * w;
    real *F = reinterpret_cast<real *>(p);
    int *tbl = reinterpret_cast<int *>(p + (PLAIN ? 0 : (size_t)Np * sizeof(real)));
    uint32_t *Qb = reinterpret_cast<uint32_t *>(tbl + tblN);
    uint32_t *Mir[2] = {Qb + Nw, Qb + 2 * Nw};

    for (int i = threadIdx.x; i < Np; i += blockDim.x)
        hS[i] = a.h[i];
    __syncthreads();
    if (y >= m)
        return;

    const long long yN = (long long)y * N;
    const int n0 = (y == 0) ? m - 1 : y - 1;
    const int n1 = (y == m - 1) ? 0 : y + 1;
    const int nbr[2] = {n0, n1};
    const int posY = slicePos(y, m);
    const int delta[2] = {slicePos(n0, m) < posY ? 1 : 0, slicePos(n1, m) < posY ? 1 : 0};

    if (!PLAIN)
        for (int i = lane; i < Np; i += kWave)
            F[i] = a.F[(long long)y * Np + i];
    for (int i = lane; i < tblN; i += kWave)
        tbl[i] = 0x7fffffff;
    for (int i = lane; i < Nw; i += kWave) {
        Qb[i] = a.Qbits[(long long)y * Nw + i];
        Mir[0][i] = a.Qbits[(long long)n0 * Nw + i];
        Mir[1][i] = a.Qbits[(long long)n1 * Nw + i];
    }
    waveSync();

    const uint32_t tag = a.tag;
    int applied[2] = {0, 0};
    int k = 0;
    unsigned long long nFlips = 0, nBatches = 0, nBlocked = 0;
    int spins = 0;
    bool aborted = false;

    while (k < N) {
        /* own attempts k + lane */
        const int kk = k + lane;
        const bool inRange = kk < N;
        const uint32_t s = inRange ? a.X[yN + kk] : 0u;
        const real u = inRange ? a.U[yN + kk] : (real)0;
        int limit = N - k;

        /* neighbours: replay decided prefixes, register pending entries */
        uint32_t pendSite[2][2];
        bool pendSet[2][2];
        if (!a.sa) {
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const long long nN = (long long)nbr[nb] * N;
                const int Kk = k + delta[nb];
                int base;
                uint32_t d0, d1, s0, s1;
                int e0, e1;
                for (;;) {
                    base = applied[nb];
                    e0 = base + lane;
                    e1 = base + kWave + lane;
                    d0 = e0 < N ? loadShared32(a.D + nN + e0) : 0u;
                    d1 = e1 < N ? loadShared32(a.D + nN + e1) : 0u;
                    s0 = e0 < N ? a.X[nN + e0] : 0u;
                    s1 = e1 < N ? a.X[nN + e1] : 0u;
                    const bool dec0 = (d0 >> 1) == tag, dec1 = (d1 >> 1) == tag;
                    const unsigned long long u0 = __ballot(e0 < N && !dec0);
                    const unsigned long long u1 = __ballot(e1 < N && !dec1);
                    int winLen = N - base < 2 * kWave ? N - base : 2 * kWave;
                    int pfx = u0 ? firstLane(u0) : (u1 ? kWave + firstLane(u1) : winLen);
                    int applyEnd = base + pfx < Kk ? base + pfx : Kk;
                    if (e0 < applyEnd && (d0 & 1u))
                        atomicXor(&Mir[nb][s0 >> 5], 1u << (s0 & 31));
                    if (e1 < applyEnd && (d1 & 1u))
                        atomicXor(&Mir[nb][s1 >> 5], 1u << (s1 & 31));
                    applied[nb] = applyEnd;
                    if (pfx == 2 * kWave && applyEnd == base + 2 * kWave && applyEnd < Kk)
                        continue; /* whole window replayed, still behind: next window */
                    break;
                }
                const int A = applied[nb];
                pendSet[nb][0] = e0 >= A && e0 < N;
                pendSet[nb][1] = e1 >= A && e1 < N;
                pendSite[nb][0] = s0;
                pendSite[nb][1] = s1;
                if (pendSet[nb][0])
                    atomicMin(&tbl[s0 & tblMask], e0 - Kk);
                if (pendSet[nb][1])
                    atomicMin(&tbl[s1 & tblMask], e1 - Kk);
                /* lane i needs entries [A, Kk + i): they must lie in the window */
                const int winEnd = base + 2 * kWave < N ? base + 2 * kWave : N;
                if (winEnd < N) {
                    int lim = winEnd - Kk + 1;
                    limit = lim < limit ? lim : limit;
                }
            }
        }
        waveSync();

        /* evaluate attempt k + lane speculatively */
        const bool valid = lane < limit;
        bool blocked = false;
        if (!a.sa)
            blocked = tbl[s & tblMask] < lane;
        const uint32_t qword = Qb[s >> 5];
        const bool qpos = (qword >> (s & 31)) & 1u;
        real f;
        if constexpr (PLAIN) {
            /* reference algorithm: field = J[s] . q_y, recomputed per attempt */
            real acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
            if (valid && !blocked) {
                const long long rowOff = (long long)s * Np;
                for (int j = 0; j < Np; j += 32) {
                    const uint32_t qw = Qb[j >> 5];
#pragma unroll
                    for (int t = 0; t < 32; t += 4) {
                        real v[4];
                        JRow<real, fieldF32>::load4(a, rowOff + j + t, v);
                        acc0 += ((qw >> (t + 0)) & 1u) ? v[0] : -v[0];
                        acc1 += ((qw >> (t + 1)) & 1u) ? v[1] : -v[1];
                        acc2 += ((qw >> (t + 2)) & 1u) ? v[2] : -v[2];
                        acc3 += ((qw >> (t + 3)) & 1u) ? v[3] : -v[3];
                    }
                }
            }
            f = (acc0 + acc1) + (acc2 + acc3);
        } else {
            f = F[s];
        }
        const real q = qpos ? (real)1 : (real)-1;
        const real hx = hS[s];
        real thr;
        if (!a.sa) {
            const real nv0 = ((Mir[0][s >> 5] >> (s & 31)) & 1u) ? (real)1 : (real)-1;
            const real nv1 = ((Mir[1][s >> 5] >> (s & 31)) & 1u) ? (real)1 : (real)-1;
            /* CPUDenseGraphAnnealer.cpp:258-262, same operation order */
            real dE = a.twoDivM * q * (hx + (real)2. * f);
            dE -= q * (nv0 + nv1) * a.coef;
            thr = (dE < (real)0.) ? (real)1. : Ops<real>::expo(-dE * a.beta);
        } else {
            /* :348-349 */
            real dE = (real)2. * q * (hx + (real)2. * f);
            thr = (dE < (real)0.) ? (real)1. : Ops<real>::expo(-dE * a.invKT);
        }
        const bool acc = thr > u;
        const unsigned long long stopMask = __ballot(!valid || blocked);
        const unsigned long long accMask = __ballot(valid && !blocked && acc);
        const int firstStop = firstLane(stopMask);
        const int firstAcc = firstLane(accMask);
        int c, flipLane;
        if (firstAcc < firstStop) {
            c = firstAcc + 1;
            flipLane = firstAcc;
        } else {
            c = firstStop;
            flipLane = -1;
        }
        waveSync();

        /* clear the pending-site table */
        if (!a.sa) {
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                if (pendSet[nb][0])
                    tbl[pendSite[nb][0] & tblMask] = 0x7fffffff;
                if (pendSet[nb][1])
                    tbl[pendSite[nb][1] & tblMask] = 0x7fffffff;
            }
        }

        if (c == 0) {
            /* lane 0 waits on a neighbour's serially earlier attempt */
            ++nBlocked;
            if (++spins > a.spinLimit) {
                atomicExch(a.abortFlag, 1);
                aborted = true;
                break;
            }
            if ((spins & 63) == 0 && __hip_atomic_load(a.abortFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                aborted = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            waveSync();
            continue;
        }
        spins = 0;
        ++nBatches;

        /* publish the committed decisions */
        if (lane < c)
            storeShared32(a.D + yN + k + lane, (tag << 1) | (lane == flipLane ? 1u : 0u));

        if (flipLane >= 0) {
            const uint32_t x = __shfl(s, flipLane);
            const bool qx = __shfl((int)qpos, flipLane) != 0;
            ++nFlips;
            if (lane == 0)
                atomicXor(&Qb[x >> 5], 1u << (x & 31));
            if constexpr (!PLAIN) {
                /* F[y][:] += -2 q_x J[:][x] = -2 q_x J[x][:]  (J symmetric) */
                const real d = qx ? (real)-2. : (real)2.;
                const long long rowOff = (long long)x * Np;
                for (int j = lane * 4; j < Np; j += kWave * 4) {
                    real v[4];
                    JRow<real, FMT>::load4(a, rowOff + j, v);
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        F[j + t] = F[j + t] + d * v[t];
                }
            }
        }
        waveSync();
        k += c;
    }

    if (!aborted) {
        if (!PLAIN)
            for (int i = lane; i < Np; i += kWave)
                a.F[(long long)y * Np + i] = F[i];
        for (int i = lane; i < Nw; i += kWave)
            a.Qbits[(long long)y * Nw + i] = Qb[i];
    }
    if (lane == 0 && a.counters) {
        atomicAdd(&a.counters[0], nFlips);
        atomicAdd(&a.counters[1], nBatches);
        atomicAdd(&a.counters[2], nBlocked);
    }
}

}  // namespace

size_t denseSweepLdsBytes(int Np, int slicesPerBlock, size_t realSize, bool plain) {
    int tblN = pendingTableSize(Np);
    int Nw = Np / 32;
    size_t perWave = (plain ? 0 : (size_t)Np * realSize) + (size_t)tblN * 4 + (size_t)3 * Nw * 4;
    return (size_t)Np * realSize + perWave * slicesPerBlock;
}

template <class real>
void denseSweep(const SweepArgs<real> &args, bool plain, int fieldFormat, int slicesPerBlock, hipStream_t stream) {
    SQ_THROW_IF(args.Np % 256 != 0, "padded N must be a multiple of 256");
    SQ_THROW_IF(slicesPerBlock < 1 || slicesPerBlock > 16, "slices per block out of range");
    SQ_THROW_IF(!plain && fieldFormat == fieldBF16x2 && sizeof(real) != 4, "bf16x2 fields need float");
    int nBlocks = divru(args.m, slicesPerBlock);
    size_t lds = denseSweepLdsBytes(args.Np, slicesPerBlock, sizeof(real), plain);
    int xcdRemap = (nBlocks % 8 == 0 && nBlocks >= 16) ? 1 : 0;

    void *fn;
    if (plain)
        fn = (void *)denseSweepKernel<real, true, fieldF32>;
    else if (fieldFormat == fieldBF16x2)
        fn = (void *)denseSweepKernel<real, false, fieldBF16x2>;
    else
        fn = (void *)denseSweepKernel<real, false, fieldF32>;
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    SweepArgs<real> a = args;
    void *kargs[] = {(void *)&a, (void *)&slicesPerBlock, (void *)&xcdRemap};
    HIP_CHECK(hipLaunchCooperativeKernel(fn, dim3(nBlocks), dim3(slicesPerBlock * kWave), kargs, (unsigned)lds,
                                         stream));
}

template void denseSweep<float>(const SweepArgs<float> &, bool, int, int, hipStream_t);
template void denseSweep<double>(const SweepArgs<double> &, bool, int, int, hipStream_t);

}  // namespace sqaod_rocm
