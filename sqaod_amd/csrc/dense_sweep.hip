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
 * of a hang.  All waves must be co-resident (see denseSweep's launch).
 */
#include "common.h"
#include "kernels.h"
#include "sq_expf.h"

#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

namespace sqaod_rocm {

namespace {

constexpr int kWave = 64;
constexpr int kTblMax = 4096; /* entries of the pending-site table per slice */
constexpr int kDotGroup = 8; /* plain sweep: attempts whose dots are formed together */
constexpr int kMaxSlicesPerBlock = 4; /* one wave per SIMD, so a wave may hold up to 256 VGPRs */
constexpr size_t kExpTabBytes = 32 * sizeof(uint64_t); /* LDS copy of the exp table */

/* Optional per-phase cycle counters (build with -DSQAOD_SWEEP_CYCLES): shader
 * clock cycles spent per phase, summed into SweepArgs::counters[3..15]. */
#ifdef SQAOD_SWEEP_CYCLES
#define SQ_CYC(v) const unsigned long long v = __builtin_amdgcn_s_memtime()
#define SQ_CYC_ADD(acc, a, b) (acc) += (b) - (a)
#define SQ_CYC_WAITALL() __builtin_amdgcn_s_waitcnt(0)
#else
#define SQ_CYC(v)
#define SQ_CYC_ADD(acc, a, b)
#define SQ_CYC_WAITALL()
#endif


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

/* lane l's float (l wave-uniform) */
__device__ __forceinline__ float readLaneF(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

/* A wave-uniform int, marked so: control values the compiler cannot prove
 * uniform (they pass through loops with ballot-dependent exits) would
 * otherwise live in VGPRs and turn every branch on them into exec masking. */
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

/* LDS DMA issued from inline assembly: lane i's 4 (16) bytes from src land at
 * ldsDst + 4i (16i).  The compiler sees neither the op nor its LDS write, so
 * it adds no memory waits for it (for the builtin form it waits for every
 * outstanding DMA before any later LDS access, which serialises the batch
 * loop); every read of a DMA destination follows an explicit vmWait instead.
 * Ops the compiler does not count only make its own waits stricter (vmcnt
 * retires in issue order), never weaker.  M0 carries the LDS address; nothing
 * else in the sweep kernels uses M0.  kCoherent: the cache policy of a relaxed
 * atomic load at agent scope (decisions published by other waves). */
/* the M0 clobber below is deliberate (clang warns about any M0 clobber) */
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ uint32_t ldsAddress(const void *p) {
    return __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p);
}
template <bool kCoherent>
__device__ __forceinline__ void dmaDword(const void *src, uint32_t lds) {
    const uint32_t m = __builtin_amdgcn_readfirstlane(lds);
    if constexpr (kCoherent)
        asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dword %0, off sc1" ::"v"(src), "s"(m) : "memory", "m0");
    else
        asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dword %0, off" ::"v"(src), "s"(m) : "memory", "m0");
}
__device__ __forceinline__ void dmaDwordx4(const void *src, uint32_t lds) {
    const uint32_t m = __builtin_amdgcn_readfirstlane(lds);
    asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(m) : "memory", "m0");
}
/* n consecutive 1 KiB DMA ops (16 B per lane): op i reads src + i KiB (src:
 * this lane's address for op 0) and lands at LDS byte address lds + i KiB.
 * Groups of four set M0 once and step it by 1 KiB (a wait state after each
 * M0 write before the DMA reads it). */
__device__ __forceinline__ void dmaRunx4(const unsigned char *src, uint32_t lds, int n) {
    int i = 0;
    for (; i + 4 <= n; i += 4) { /* n is wave-uniform: a scalar loop */
        const uint32_t m = __builtin_amdgcn_readfirstlane(lds + (uint32_t)i * 1024u);
        const unsigned char *p = src + (size_t)i * 1024;
        asm volatile("s_mov_b32 m0, %4\n\ts_nop 0\n\t"
                     "global_load_lds_dwordx4 %0, off\n\t"
                     "s_add_u32 m0, m0, 0x400\n\ts_nop 0\n\t"
                     "global_load_lds_dwordx4 %1, off\n\t"
                     "s_add_u32 m0, m0, 0x400\n\ts_nop 0\n\t"
                     "global_load_lds_dwordx4 %2, off\n\t"
                     "s_add_u32 m0, m0, 0x400\n\ts_nop 0\n\t"
                     "global_load_lds_dwordx4 %3, off" ::"v"(p),
                     "v"(p + 1024), "v"(p + 2048), "v"(p + 3072), "s"(m)
                     : "memory", "m0", "scc");
    }
    for (; i < n; ++i)
        dmaDwordx4(src + (size_t)i * 1024, lds + (uint32_t)i * 1024u);
}
/* A neighbour's two 128-entry windows when all their entries lie in its row:
 * the decisions (coherent loads) at LDS byte address lds and the sites 1 KiB
 * above it, 256 B per op, four ops on one M0 chain.  d, x: this lane's entry
 * 0 address in each stream. */
__device__ __forceinline__ void dmaNeighbourWindows(const uint32_t *d, const uint32_t *x, uint32_t lds) {
    const uint32_t m = __builtin_amdgcn_readfirstlane(lds);
    asm volatile("s_mov_b32 m0, %4\n\ts_nop 0\n\t"
                 "global_load_lds_dword %0, off sc1\n\t"
                 "s_add_u32 m0, m0, 0x100\n\ts_nop 0\n\t"
                 "global_load_lds_dword %1, off sc1\n\t"
                 "s_add_u32 m0, m0, 0x300\n\ts_nop 0\n\t"
                 "global_load_lds_dword %2, off\n\t"
                 "s_add_u32 m0, m0, 0x100\n\ts_nop 0\n\t"
                 "global_load_lds_dword %3, off" ::"v"(d),
                 "v"(d + 64), "v"(x), "v"(x + 64), "s"(m)
                 : "memory", "m0", "scc");
}
#pragma clang diagnostic pop

/* wait until at most C vector-memory ops are outstanding */
template <int C>
__device__ __forceinline__ void vmWaitImm() {
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(C) : "memory");
}
template <int LO, int HI>
__device__ __forceinline__ void vmWaitRange(int c) {
    if constexpr (LO == HI) {
        vmWaitImm<LO>();
    } else {
        constexpr int MID = (LO + HI + 1) / 2;
        if (c >= MID)
            vmWaitRange<MID, HI>(c);
        else
            vmWaitRange<LO, MID - 1>(c);
    }
}
/* vmcnt(c) for a wave-uniform c (counts above 63 wait for 63) */
__device__ __forceinline__ void vmWait(int c) {
    c = __builtin_amdgcn_readfirstlane(c);
    if (c == 1) /* the common case: only a decision store since */
        vmWaitImm<1>();
    else
        vmWaitRange<0, 63>(c < 63 ? c : 63);
}
/* this wave's LDS reads have returned (before a DMA overwrites what they read) */
__device__ __forceinline__ void ldsWait() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

/* Copy kEntries entries of ES bytes, from entry e0 of src, into LDS at lds by
 * DMA, skipping entries >= limit (their LDS words keep old contents). */
template <int ES, bool kCoherent, int kEntries>
__device__ __forceinline__ void dmaEntries(const void *src, uint32_t lds, int e0, int limit) {
    const int lane = threadIdx.x & (kWave - 1);
    constexpr int kOps = kEntries * ES / (4 * kWave);
    const uint32_t *s = reinterpret_cast<const uint32_t *>(src) + lane;
    if (e0 + kEntries <= limit) {
#pragma unroll
        for (int j = 0; j < kOps; ++j)
            dmaDword<kCoherent>(s + j * kWave, lds + j * kWave * 4);
    } else {
#pragma unroll
        for (int j = 0; j < kOps; ++j)
            if (e0 + (j * kWave + lane) * 4 / ES < limit)
                dmaDword<kCoherent>(s + j * kWave, lds + j * kWave * 4);
    }
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

/* Pending-site table codes: the serially first undecided neighbour attempt
 * registered at a slot, as ((e - Kk + N) << 1) | nb (unique per attempt),
 * kPendMulti when more than one attempt hashed there, kPendEmpty when none. */
constexpr int kPendEmpty = 0x7fffffff;
constexpr int kPendMulti = 1 << 30;
__device__ __forceinline__ int pendCode(int rel, int N, int nb) { return ((rel + N) << 1) | nb; }
__device__ __forceinline__ int pendRel(int code, int N) { return ((code & (kPendMulti - 1)) >> 1) - N; }

__device__ __forceinline__ int slicePos(int y, int m) {
    int nEven = (m + 1) / 2;
    return (y & 1) ? nEven + (y >> 1) : (y >> 1);
}

/* 4 consecutive elements of a J row (16- or 32-byte vector load). */
template <class real>
__device__ __forceinline__ void loadJ4(const real *p, real v[4]) {
    if constexpr (sizeof(real) == 4) {
        float4 t = *reinterpret_cast<const float4 *>(p);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
        double2 t0 = *reinterpret_cast<const double2 *>(p);
        double2 t1 = *reinterpret_cast<const double2 *>(p + 2);
        v[0] = t0.x; v[1] = t0.y; v[2] = t1.x; v[3] = t1.y;
    }
}

template <class real>
__device__ __forceinline__ void ldsLoad4(const real *p, real v[4]) {
    if constexpr (sizeof(real) == 4) {
        float4 t = *reinterpret_cast<const float4 *>(p);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
        double2 t0 = *reinterpret_cast<const double2 *>(p);
        double2 t1 = *reinterpret_cast<const double2 *>(p + 2);
        v[0] = t0.x; v[1] = t0.y; v[2] = t1.x; v[3] = t1.y;
    }
}

template <class real>
__device__ __forceinline__ void ldsStore4(real *p, const real v[4]) {
    if constexpr (sizeof(real) == 4) {
        *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
        *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2 *>(p + 2) = make_double2(v[2], v[3]);
    }
}

template <class real>
struct Expo;
template <>
struct Expo<float> {
    /* the CPU annealer's expf(-dE*beta), bit-identical (sq_expf.h) */
    __device__ static float eval(float v, const uint64_t *tab) { return sq_expf_tab(v, tab); }
};
template <>
struct Expo<double> {
    __device__ static double eval(double v, const uint64_t *) { return exp(v); }
};

/* Kernel variants: local fields with rows of args.J, the plain reference
 * algorithm, and local fields with the fp8 rows of args.J8 (float only). */
enum SweepMode { kModeLocal = 0, kModePlain = 1, kModeFP8 = 2 };

/* 4 e4m3 bytes -> floats (exact), gfx950's OCP FP8 conversion */
__device__ __forceinline__ void fp8x4ToFloat(uint32_t w, float v[4]) {
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    const f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)w, false);
    const f32x2 hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)w, true);
    v[0] = lo.x; v[1] = lo.y; v[2] = hi.x; v[3] = hi.y;
}

/* Fixed part of a wave's LDS region: the two neighbour windows (decisions and
 * sites of kWin entries each), the own-stream ring (sites and uniforms of
 * kRing entries) and the fp8 row scale (a 16-byte slot). */
constexpr int kWin = 2 * kWave;
constexpr int kRing = 4 * kWave;
template <class real>
__host__ __device__ constexpr size_t sweepFixedBytes() {
    return (size_t)4 * kWin * 4 + (size_t)kRing * (4 + sizeof(real)) + 16;
}

/* Multi-flip batches (float local-field form) cover the first kMultiLanes
 * attempts of a batch; the J submatrix of their sites, element [j][i] =
 * J[s_i][s_j] for lanes i > j, is gathered one scattered dword per element,
 * so its memory traffic grows with the square of the lane count.  (32 lanes,
 * a quarter of the traffic, measured no faster than 64: the batches end
 * earlier.) */
constexpr int kMultiLanes = kWave;
constexpr size_t kSubBytes = (size_t)kMultiLanes * kWave * 4;

template <class real, int MODE>
__host__ __device__ constexpr bool multiFlipCapable() {
    return MODE == kModeLocal && sizeof(real) == 4;
}

/* Helper wave (float local-field sweep, one slice per workgroup): a second
 * row stage buffer and the command word the two waves exchange. */
constexpr size_t kHelperCtlBytes = 256;
constexpr int kHelperRounds = 4; /* the helper's row add: 256-element slices in flight */

template <class real, int MODE>
__host__ __device__ constexpr bool helperCapable() {
    return MODE == kModeLocal && sizeof(real) == 4;
}

/* Bytes of LDS one wave's slice needs (the fixed part, fields, row stage,
 * pending table, spin bits and the two mirrors), for a row length Np; with
 * multiBufs >= 2 also multiBufs - 1 more row stage buffers and the
 * submatrix of multi-flip batches; with helper (helperCapable forms) a
 * second row stage buffer and the helper's command word. */
template <class real, int MODE>
__host__ __device__ inline size_t sweepWaveBytes(int Np, int ld8, int multiBufs = 0, bool helper = false) {
    constexpr bool PLAIN = MODE == kModePlain;
    constexpr bool FP8 = MODE == kModeFP8;
    constexpr bool kDefer = !PLAIN && sizeof(real) == 4;
    const size_t stageBytes = kDefer ? (FP8 ? (size_t)ld8 : (size_t)Np * sizeof(real)) : 0;
    const size_t fieldBytes = PLAIN ? 0 : (size_t)Np * sizeof(real) + stageBytes;
    size_t b = sweepFixedBytes<real>() + fieldBytes + (size_t)pendingTableSize(Np) * 4 + (size_t)3 * (Np / 32) * 4;
    if (multiFlipCapable<real, MODE>() && multiBufs >= 2)
        b += (size_t)(multiBufs - 1) * stageBytes + kSubBytes;
    else if (helperCapable<real, MODE>() && helper)
        b += stageBytes + kHelperCtlBytes;
    return b;
}

/* Helper-wave commands (the word at the end of the slice's region): the
 * sweeping wave writes "apply buffer b with factor -+2" once the row has
 * landed in LDS (its own vmcnt wait); the helper adds the row to F, waits for
 * its LDS stores and writes kHelpIdle; kHelpExit ends it.  The word is read
 * and written with relaxed LDS atomics: a wave's LDS operations execute in
 * issue order, and the compiler barriers around them keep the field accesses
 * on their side of the exchange. */
constexpr uint32_t kHelpIdle = 0, kHelpExit = 1;
__device__ __forceinline__ uint32_t helpApply(int buf, bool negative) {
    return 2u + ((uint32_t)buf << 1) + (negative ? 1u : 0u);
}
__device__ __forceinline__ uint32_t helpLoad(const uint32_t *ctl) {
    const uint32_t v = __hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    asm volatile("" ::: "memory");
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ void helpStore(uint32_t *ctl, uint32_t v) {
    asm volatile("" ::: "memory");
    if ((threadIdx.x & (kWave - 1)) == 0)
        __hip_atomic_store(ctl, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
/* the slice region's helper command word: after the second stage buffer,
 * which follows the spin bits and mirrors */
template <class real>
__device__ __forceinline__ uint32_t *helperCtl(const SweepArgs<real> &a, unsigned char *p) {
    const size_t stageBytes = (size_t)a.Np * sizeof(real);
    unsigned char *body = p + sweepFixedBytes<real>();
    unsigned char *bits = body + (size_t)a.Np * sizeof(real) + stageBytes + (size_t)pendingTableSize(a.Np) * 4;
    unsigned char *buf1 = bits + (size_t)3 * a.Nw * 4;
    return reinterpret_cast<uint32_t *>(buf1 + stageBytes);
}

/* f + d * v for a field update: d is -+2 (fp8 rows: -+2 times the row's
 * power-of-two scale) and v a J element, so d * v is exact and one fused
 * multiply-add rounds exactly as the separate multiply and add do. */
__device__ __forceinline__ float addScaled(float f, float d, float v) { return __builtin_fmaf(d, v, f); }
__device__ __forceinline__ double addScaled(double f, double d, double v) { return __builtin_fma(d, v, f); }

/* F[b0 ..] += d * sb[b0 ..] for R consecutive 256-element slices: all their
 * LDS reads are issued before the first write, so the round trips overlap.
 * FP8: sb holds e4m3 bytes. */
template <class real, bool FP8, int R>
__device__ __forceinline__ void addStagedSlices(real *F, const real *sb, real d, int b0, int lane) {
#pragma clang fp contract(off)
    const uint8_t *sb8 = reinterpret_cast<const uint8_t *>(sb);
    (void)sb8;
    real fv[R][4], sv[R][4];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = b0 + r * kWave * 4 + lane * 4;
        ldsLoad4(F + j, fv[r]);
        if constexpr (FP8)
            fp8x4ToFloat(*reinterpret_cast<const uint32_t *>(sb8 + j), sv[r]);
        else
            ldsLoad4(sb + j, sv[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = b0 + r * kWave * 4 + lane * 4;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            fv[r][t] = addScaled(fv[r][t], d, sv[r][t]);
        ldsStore4(F + j, fv[r]);
    }
}

/* F += d * (staged row sb): rounds of kR 256-element slices, then single
 * slices for the rest (Np is a multiple of 256).  The rounds carry no
 * per-slice bounds checks, so their reads go out back to back. */
template <class real, bool FP8, int kR = 4>
__device__ __forceinline__ void addStagedRow(real *F, const real *sb, real d, int Np, int lane) {
    int b0 = 0;
    /* b0 and the bounds are wave-uniform: scalar loops */
    for (; b0 + kR * kWave * 4 <= Np; b0 += kR * kWave * 4)
        addStagedSlices<real, FP8, kR>(F, sb, d, b0, lane);
    for (; b0 < Np; b0 += kWave * 4)
        addStagedSlices<real, FP8, 1>(F, sb, d, b0, lane);
}

/* F += d[0] * row 0, then d[1] * row 1, ... (NR rows, in that order per
 * element), reading and writing F once for all of them: rounds of RR
 * 256-element slices, then single slices for the rest. */
template <class real, int NR>
__device__ __forceinline__ void addStagedRows(real *F, const real *const *sb, const real *d, int Np, int lane) {
#pragma clang fp contract(off)
    auto slices = [&](int b0, auto rTag) {
        constexpr int R = decltype(rTag)::value;
        real fv[R][4], sv[R][NR][4];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = b0 + r * kWave * 4 + lane * 4;
            ldsLoad4(F + j, fv[r]);
#pragma unroll
            for (int q = 0; q < NR; ++q)
                ldsLoad4(sb[q] + j, sv[r][q]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = b0 + r * kWave * 4 + lane * 4;
#pragma unroll
            for (int q = 0; q < NR; ++q)
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    fv[r][t] = addScaled(fv[r][t], d[q], sv[r][q][t]);
            ldsStore4(F + j, fv[r]);
        }
    };
    constexpr int RR = NR <= 2 ? 4 : 2; /* slices per round: (1 + NR) RR reads in flight */
    int b0 = 0;
    for (; b0 + RR * kWave * 4 <= Np; b0 += RR * kWave * 4)
        slices(b0, std::integral_constant<int, RR>());
    for (; b0 < Np; b0 += kWave * 4)
        slices(b0, std::integral_constant<int, 1>());
}

/* The helper wave of a slice (float local-field sweep): waits for commands
 * from the sweeping wave and applies the rows it hands over, F += d * row,
 * with the same per-element operation the sweeping wave would use. */
template <class real>
__device__ void helperLoop(const SweepArgs<real> &a, unsigned char *p) {
    const int Np = a.Np;
    const int lane = threadIdx.x & (kWave - 1);
    real *F = reinterpret_cast<real *>(p + sweepFixedBytes<real>());
    uint32_t *ctl = helperCtl(a, p);
    real *const buf[2] = {F + Np, reinterpret_cast<real *>(reinterpret_cast<unsigned char *>(ctl) - (size_t)Np * sizeof(real))};
    for (;;) {
        const uint32_t c = helpLoad(ctl);
        if (c == kHelpIdle) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        if (c == kHelpExit)
            break;
        const uint32_t cmd = c - 2u;
        addStagedRow<real, false, kHelperRounds>(F, buf[cmd >> 1], (cmd & 1u) ? (real)-2. : (real)2., Np, lane);
        ldsWait(); /* the field stores are done */
        helpStore(ctl, kHelpIdle);
    }
}

/* One wave runs the whole sweep of slice y: N attempts, with its LDS region
 * p (sweepWaveBytes), the exp table and h already in LDS.  MULTI compiles the
 * multi-flip batches in (float local-field form only); without it the batch
 * ends at its first accepted flip. */
template <class real, int MODE, bool MULTI = false, bool HELPER = false>
__device__ __forceinline__ void sweepSlice(const SweepArgs<real> &a, const int y, const uint64_t *expTab,
                                           const real *hS, unsigned char *p) {
#pragma clang fp contract(off)
    constexpr bool PLAIN = MODE == kModePlain;
    constexpr bool FP8 = MODE == kModeFP8;
    static_assert(!FP8 || sizeof(real) == 4, "fp8 rows are for float solvers");
    const int N = a.N, Np = a.Np, m = a.m, Nw = a.Nw;
    const int lane = threadIdx.x & (kWave - 1);
    const int tblN = pendingTableSize(Np);
    const int tblMask = tblN - 1;
    constexpr bool kDefer = !PLAIN && sizeof(real) == 4; /* deferred row updates, staged in LDS */
    const size_t stageBytes = kDefer ? (FP8 ? (size_t)a.ld8 : (size_t)Np * sizeof(real)) : 0;
    const size_t fieldBytes = PLAIN ? 0 : (size_t)Np * sizeof(real) + stageBytes;
    uint32_t *const winD = reinterpret_cast<uint32_t *>(p); /* [nb][kWin] neighbour decisions */
    uint32_t *const winS = winD + 2 * kWin;                  /* [nb][kWin] neighbour sites */
    uint32_t *const ringX = winS + 2 * kWin;                 /* own sites, entry e at e % kRing */
    real *const ringU = reinterpret_cast<real *>(ringX + kRing);
    float *const scaleSlot = reinterpret_cast<float *>(ringU + kRing); /* FP8: the last flip's row scale */
    unsigned char *const body = p + sweepFixedBytes<real>();
    real *F = reinterpret_cast<real *>(body);
    real *stage = F + Np; /* kDefer: the J row of the last flip, as loaded */
    const uint8_t *stage8 = reinterpret_cast<const uint8_t *>(stage); /* FP8: the e4m3 row */
    int *tbl = reinterpret_cast<int *>(body + fieldBytes);
    uint32_t *Qb = reinterpret_cast<uint32_t *>(tbl + tblN);
    uint32_t *Mir[2] = {Qb + Nw, Qb + 2 * Nw};
    /* LDS byte addresses of the DMA destinations (wave-uniform) */
    const uint32_t winDL = ldsAddress(p), winSL = winDL + 2 * kWin * 4, ringXL = winSL + 2 * kWin * 4;
    const uint32_t ringUL = ringXL + kRing * 4, scaleL = ringUL + kRing * (uint32_t)sizeof(real);
    const uint32_t stageL = winDL + (uint32_t)(sweepFixedBytes<real>() + (size_t)Np * sizeof(real));
    (void)scaleL;
    (void)stageL;
    /* Multi-flip batches: buffers 1 .. nBufs-1 of row stage after the
     * mirrors, then the submatrix.  A batch in multi mode gathers, for its 64
     * sites, sub[j][i] = J[s_i][s_j] (i > j) by DMA while the neighbour phase
     * runs; its lanes then decide in order, each accepted flip correcting the
     * fields of the later lanes with that flip's column of the submatrix,
     * and the commit applies the batch's flipped rows to F through the stage
     * buffers (the last one deferred like a single flip's). */
    constexpr bool kMultiCap = MULTI && multiFlipCapable<real, MODE>();
    const int nBufs = kMultiCap ? uni(a.multiBufs) : 1;
    unsigned char *const multiBase = reinterpret_cast<unsigned char *>(Mir[1] + Nw);
    real *const sub = reinterpret_cast<real *>(multiBase + (size_t)(nBufs - 1) * stageBytes);
    const uint32_t multiBaseL = ldsAddress(multiBase);
    const uint32_t subL = multiBaseL + (uint32_t)((size_t)(nBufs - 1) * stageBytes);
    (void)sub;
    (void)subL;
    /* row stage buffer b: buffer 0 is `stage`, the others follow the mirrors */
    auto stageBuf = [&](int b) -> real * {
        return b == 0 ? stage : reinterpret_cast<real *>(multiBase + (size_t)(b - 1) * stageBytes);
    };
    auto stageBufL = [&](int b) -> uint32_t {
        return b == 0 ? stageL : multiBaseL + (uint32_t)((size_t)(b - 1) * stageBytes);
    };
    (void)stageBuf;
    (void)stageBufL;
    /* Helper wave: the pending row is handed to the slice's second wave at
     * the commit instead of being applied there, and the next row streams
     * into the other of stage buffers 0 and 1.  The handed row must be in F
     * before the next evaluation reads F (helpBusy: wait for kHelpIdle). */
    constexpr bool kHelp = HELPER && kDefer && !FP8 && !kMultiCap && helperCapable<real, MODE>();
    uint32_t *const ctl = reinterpret_cast<uint32_t *>(multiBase + stageBytes);
    bool helpBusy = false;
    (void)ctl;
    (void)helpBusy;

    const long long yN = (long long)y * N;
    const int n0 = (y == 0) ? m - 1 : y - 1;
    const int n1 = (y == m - 1) ? 0 : y + 1;
    const int nbr[2] = {n0, n1};
    const int posY = slicePos(y, m);
    const int delta[2] = {slicePos(n0, m) < posY ? 1 : 0, slicePos(n1, m) < posY ? 1 : 0};

    /* start-up state by DMA, all in flight together (the first batch waits
     * for it): fields (1 KiB per op), spin bits and the two mirrors */
    if (!PLAIN) {
        const unsigned char *src = reinterpret_cast<const unsigned char *>(a.F + (long long)y * Np);
        const uint32_t FL = winDL + (uint32_t)sweepFixedBytes<real>();
        for (int j0 = 0; j0 < Np * (int)sizeof(real); j0 += kWave * 16)
            dmaDwordx4(src + j0 + lane * 16, FL + j0);
    }
    {
        const uint32_t QbL = ldsAddress(Qb);
        for (int b = 0; b < Nw; b += kWave) {
            dmaEntries<4, false, kWave>(a.Qbits + (long long)y * Nw + b, QbL + b * 4, b, Nw);
            dmaEntries<4, false, kWave>(a.Qbits + (long long)n0 * Nw + b, QbL + (Nw + b) * 4, b, Nw);
            dmaEntries<4, false, kWave>(a.Qbits + (long long)n1 * Nw + b, QbL + (2 * Nw + b) * 4, b, Nw);
        }
    }
    for (int b = 0; b < tblN; b += kWave)
        tbl[b + lane] = kPendEmpty;

    const uint32_t tag = a.tag;
    int applied[2] = {0, 0};
    int k = 0;
    uint32_t nFlips = 0, nBatches = 0, nBlocked = 0; /* per wave and sweep: at most N each */
    int spins = 0;
    bool aborted = false;
    /* kDefer: the J row of the last committed flip is loaded straight into
     * `stage` (global_load_lds) and added to F at the next commit; meanwhile
     * evaluations add rowD * stage[s] to F[s] themselves -- the same float
     * operation as the LDS update, so fields are bit-identical.  FP8: the row
     * is e4m3 and rowD = -+2 2^a_x, so rowD * byte = -+2 J'[x][s] exactly, the
     * float the fp32 form adds for the same J'. */
    bool rowPending = false;
    real rowD = 0;
    int curStage = 0; /* the stage buffer holding the pending row */
    /* multi-flip kernel: up to nBufs deferred rows, row p in stage buffer p
     * with factor pendD[p], added to a field in row order (the serial order
     * of their flips) by the evaluations and, in one pass over F, by the next
     * commit */
    constexpr int kMaxPend = 4;
    int nPend = 0;
    real pendD[kMaxPend] = {0, 0, 0, 0};
    (void)nPend;
    (void)pendD;
    /* F += d * (row in stage buffer b) */
    auto applyRow = [&](int b, real d) { addStagedRow<real, FP8>(F, stageBuf(b), d, Np, lane); };
    auto applyPendingRow = [&]() {
        applyRow(curStage, rowD);
        rowPending = false;
    };
    /* F += pendD[0] * row 0, then pendD[1] * row 1, ... per element (the
     * serial order), reading F once for all the deferred rows */
    auto applyPendingRows = [&]() {
        if (nPend == 1) { /* one row: the deeper single-row rounds */
            applyRow(0, pendD[0]);
            nPend = 0;
            return;
        }
        const real *const sbs[kMaxPend] = {stageBuf(0), stageBuf(1), stageBuf(2), stageBuf(3)};
        if (nPend == 2)
            addStagedRows<real, 2>(F, sbs, pendD, Np, lane);
        else if (nPend == 3)
            addStagedRows<real, 3>(F, sbs, pendD, Np, lane);
        else
            addStagedRows<real, 4>(F, sbs, pendD, Np, lane);
        nPend = 0;
    };
    (void)applyPendingRows;

    /* Memory traffic of the batch loop.  What the loop reads from global
     * memory arrives in LDS by DMA (dmaDword), so its memory waits are these,
     * in issue order:
     *   neighbour phase: the prefetch group -- both neighbour windows for the
     *     next batch and at most one new quarter of the own-stream ring;
     *   commit: the decision store (1 op), and after a flip the row and, for
     *     fp8, its scale (rowOps ops).
     * A batch waits for the last prefetch group with vmcnt(ops issued since),
     * which leaves the decision store and the row in flight; the row is
     * waited for where it is first read (rowCorr). */
    const int rowOps = kDefer ? (FP8 ? a.ld8 / (kWave * 16) + 1 : Np * (int)sizeof(real) / (kWave * 16)) : 0;
    int afterGroup = 0; /* counted ops issued since the last prefetch group */
    real rowSign = 0;   /* FP8: -+2 of the staged row; rowD = rowSign * its scale */
    (void)rowSign;

    /* Own attempt stream: a ring of kRing entries (entry e at e % kRing)
     * filled a quarter (64 entries) at a time, up to two quarters beyond the
     * one holding k; a batch reads entries k .. k + 63. */
    int ringNext = 0; /* quarters [0, ringNext) have been requested */
    auto requestQuarter = [&]() {
        const int e0 = ringNext * kWave;
        const int slot = e0 & (kRing - 1);
        dmaEntries<4, false, kWave>(a.X + yN + e0, ringXL + slot * 4, e0, N);
        dmaEntries<(int)sizeof(real), false, kWave>(a.U + yN + e0, ringUL + slot * (uint32_t)sizeof(real), e0, N);
        ++ringNext;
    };

    /* Neighbour windows: decisions and sites of entries [applied, applied+128)
     * in LDS, requested one batch ahead. */
    uint32_t ns[2][2];
    int nbase[2] = {0, 0};
    auto loadNeighbour = [&](int nb) {
        const long long nN = (long long)nbr[nb] * N;
        const int base = uni(applied[nb]);
        nbase[nb] = base;
        if (base + kWin <= N) {
            static_assert(kWin == 128, "two 64-lane ops per window");
            dmaNeighbourWindows(a.D + nN + base + lane, a.X + nN + base + lane, winDL + nb * kWin * 4);
        } else {
            dmaEntries<4, true, kWin>(a.D + nN + base, winDL + nb * kWin * 4, base, N);
            dmaEntries<4, false, kWin>(a.X + nN + base, winSL + nb * kWin * 4, base, N);
        }
    };
    /* waitReads: wait for this wave's outstanding LDS reads first.  The batch
     * loop passes false: by then it has used every window entry it read (the
     * ballots and the registration addresses), so those reads have returned,
     * and the ring quarter requested is not one the batch reads (it reads
     * quarters k/64 and k/64+1, the request is at most k/64+2 of 4).  Only the
     * staged-row read of rowCorr may still be in flight, and no DMA here
     * writes the staging buffer. */
    auto prefetch = [&](bool waitReads) {
        if (waitReads)
            ldsWait();
        if (!a.sa) {
            loadNeighbour(0);
            loadNeighbour(1);
        }
        while (ringNext < (k >> 6) + 3 && ringNext * kWave < N)
            requestQuarter();
        afterGroup = 0;
    };
    prefetch(true);
    vmWaitImm<0>();

    /* load wait, neighbours, evaluate, commit, row, blocked; within neighbours:
     * replay and registration, row wait, prefetch issue; within the row:
     * the pending row's application, the next row's DMA issue; within the
     * evaluation: the wait for the helper wave; multi-flip commits: the
     * waits for rows applied at once */
    unsigned long long cyc[13] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    (void)cyc;
    /* multi mode for the next batch: chosen from how many flips the last
     * batch committed (several, or one within its first lanes) */
    bool multi = false;
    while (k < N) {
        SQ_CYC(c0);
        SQ_CYC_WAITALL();
        vmWait(afterGroup); /* the last prefetch group has landed */
        SQ_CYC(c1);
        SQ_CYC_ADD(cyc[0], c0, c1);
        int limit = uni(N - k);
        uint32_t s = ringX[(k + lane) & (kRing - 1)];
        const real u = ringU[(k + lane) & (kRing - 1)];
        if (k + lane >= N)
            s = 0u;
        if constexpr (kMultiCap) {
            if (multi) {
                /* sub[j][i] = J[s_i][s_j] for lanes j < i < kMultiLanes:
                 * kMultiLanes - 1 DMA ops of scattered dwords, landing while
                 * the neighbour phase runs (waited for with the staged row
                 * below) */
                ldsWait(); /* this wave's reads of sub (last batch) are done */
                const real *myRow = a.J + (long long)s * Np;
                for (int j = 0; j < kMultiLanes - 1; ++j) {
                    const uint32_t sj = (uint32_t)__builtin_amdgcn_readlane((int)s, j);
                    if (lane > j && lane < kMultiLanes)
                        dmaDword<false>(myRow + sj, subL + (uint32_t)j * (kWave * 4));
                }
            }
        }
        uint32_t nd[2][2]; /* the windows, read together */
        auto readWindow = [&](int nb) {
            nd[nb][0] = winD[nb * kWin + lane];
            nd[nb][1] = winD[nb * kWin + kWave + lane];
            ns[nb][0] = winS[nb * kWin + lane];
            ns[nb][1] = winS[nb * kWin + kWave + lane];
        };
        if (!a.sa) {
            readWindow(0);
            readWindow(1);
        }

        /* neighbours: replay decided prefixes, register pending entries */
        bool pendSet[2][2] = {{false, false}, {false, false}};
        bool pendShared[2][2] = {{false, false}, {false, false}}; /* the slot already held an entry */
        if (!a.sa) {
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const long long nN = (long long)nbr[nb] * N;
                const int Kk = k + delta[nb];
                for (;;) {
                    const int base = uni(nbase[nb]);
                    const int e0 = base + lane, e1 = base + kWave + lane;
                    const uint32_t d0 = nd[nb][0], d1 = nd[nb][1];
                    const bool dec0 = (d0 >> 1) == tag, dec1 = (d1 >> 1) == tag;
                    const unsigned long long u0 = __ballot(e0 < N && !dec0);
                    const unsigned long long u1 = __ballot(e1 < N && !dec1);
                    const int winLen = N - base < 2 * kWave ? N - base : 2 * kWave;
                    const int pfx = uni(u0 ? firstLane(u0) : (u1 ? kWave + firstLane(u1) : winLen));
                    const int from = uni(applied[nb]);
                    const int applyEnd = uni(base + pfx < Kk ? base + pfx : Kk);
                    const bool x0 = e0 >= from && e0 < applyEnd && (d0 & 1u);
                    const bool x1 = e1 >= from && e1 < applyEnd && (d1 & 1u);
                    if (__ballot(x0 || x1)) {
                        if (x0)
                            atomicXor(&Mir[nb][ns[nb][0] >> 5], 1u << (ns[nb][0] & 31));
                        if (x1)
                            atomicXor(&Mir[nb][ns[nb][1] >> 5], 1u << (ns[nb][1] & 31));
                    }
                    applied[nb] = uni(applyEnd > from ? applyEnd : from);
                    if (pfx == 2 * kWave && applyEnd == base + 2 * kWave && applyEnd < Kk) {
                        /* whole window replayed and still behind: fetch the next one */
                        ldsWait();
                        loadNeighbour(nb);
                        vmWaitImm<0>();
                        readWindow(nb);
                        (void)nN;
                        continue;
                    }
                    break;
                }
                const int A = uni(applied[nb]);
                const int base = uni(nbase[nb]);
                const int e0 = base + lane, e1 = base + kWave + lane;
                /* an entry matters to lane i only if e - Kk < i <= 63, and only
                 * if it may flip: a decided rejection leaves the spin as the
                 * mirror has it.  The slot keeps the serially first entry,
                 * encoded uniquely per (e, nb). */
                const bool keep0 = (nd[nb][0] >> 1) != tag || (nd[nb][0] & 1u);
                const bool keep1 = (nd[nb][1] >> 1) != tag || (nd[nb][1] & 1u);
                pendSet[nb][0] = keep0 && e0 >= A && e0 < N && e0 - Kk < kWave - 1;
                pendSet[nb][1] = keep1 && e1 >= A && e1 < N && e1 - Kk < kWave - 1;
                /* every lane issues the atomic (no exec masking); the others
                 * min with kPendEmpty, which leaves the slot as it is */
                const int old0 = atomicMin(&tbl[ns[nb][0] & tblMask], pendSet[nb][0] ? pendCode(e0 - Kk, N, nb) : kPendEmpty);
                const int old1 = atomicMin(&tbl[ns[nb][1] & tblMask], pendSet[nb][1] ? pendCode(e1 - Kk, N, nb) : kPendEmpty);
                pendShared[nb][0] = pendSet[nb][0] && old0 != kPendEmpty;
                pendShared[nb][1] = pendSet[nb][1] && old1 != kPendEmpty;
                /* lane i needs entries [A, Kk + i): they must lie in the window */
                const int winEnd = base + 2 * kWave < N ? base + 2 * kWave : N;
                if (winEnd < N) {
                    int lim = winEnd - Kk + 1;
                    limit = uni(lim < limit ? lim : limit);
                }
            }
        }
        SQ_CYC(n1);
        waveSync();
        /* a slot written by more than one entry is marked (every registration
         * after the first found it occupied): lanes meeting it wait instead
         * of deciding under both outcomes of one entry */
        int curBase[2] = {nbase[0], nbase[1]};
        if (!a.sa) {
            const bool anyShared = pendShared[0][0] || pendShared[0][1] || pendShared[1][0] || pendShared[1][1];
            if (__ballot(anyShared)) {
#pragma unroll
                for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                    for (int h = 0; h < 2; ++h)
                        if (pendShared[nb][h])
                            atomicOr(&tbl[ns[nb][h] & tblMask], kPendMulti);
                waveSync();
            }
        }
        SQ_CYC(n2);
        /* the staged row has had the whole neighbour phase to arrive: wait for
         * it before the next windows are requested (vmcnt counts in order) */
        real rowCorr = 0;
        real pendCorr[kMaxPend] = {0, 0, 0, 0};
        (void)pendCorr;
        if constexpr (kMultiCap) {
            if (nPend > 0 || multi) {
                vmWaitImm<0>();
#pragma unroll
                for (int q = 0; q < kMaxPend; ++q)
                    if (q < nPend)
                        pendCorr[q] = stageBuf(q)[s];
            }
        } else if constexpr (kDefer) {
            if (rowPending || multi) {
                vmWaitImm<0>();
                if (rowPending) {
                    if constexpr (FP8) {
                        rowD = rowSign * scaleSlot[0];
                        float v4[4];
                        fp8x4ToFloat(stage8[s], v4);
                        rowCorr = v4[0];
                    } else {
                        rowCorr = stageBuf(curStage)[s];
                    }
                }
            }
        }
        /* the next batch's neighbour windows start where this batch's replay
         * ended: request them now, so the DMA overlaps evaluation and commit */
        bool pendSetNow[2][2] = {{pendSet[0][0], pendSet[0][1]}, {pendSet[1][0], pendSet[1][1]}};
        uint32_t pendSite[2][2] = {{ns[0][0], ns[0][1]}, {ns[1][0], ns[1][1]}};
        SQ_CYC(n3);
        prefetch(false);
        SQ_CYC(c2);
        SQ_CYC_ADD(cyc[1], c1, c2);
        SQ_CYC_ADD(cyc[6], c1, n1);
        SQ_CYC_ADD(cyc[7], n2, n3);
        SQ_CYC_ADD(cyc[8], n3, c2);

        /* evaluate attempt k + lane speculatively */
        const bool valid = lane < limit;
        bool blocked = false;
        int hypNb = -1; /* one undecided neighbour attempt precedes this one at this slot */
        if (!a.sa) {
            const int code = tbl[s & tblMask];
            if (code != kPendEmpty && pendRel(code, N) < lane) {
                if (code & kPendMulti)
                    blocked = true;
                else
                    hypNb = code & 1;
            }
        }
        const bool blockedMulti = blocked; /* a slot of several pending entries: waits in any case */
        const uint32_t qword = Qb[s >> 5];
        const bool qpos = (qword >> (s & 31)) & 1u;
        real q = qpos ? (real)1 : (real)-1; /* multi mode: negated after an earlier flip at this site */
        const real hx = hS[s];
        real nsum = 0, nsumAlt = 0;
        if (!a.sa) {
            const real nv0 = ((Mir[0][s >> 5] >> (s & 31)) & 1u) ? (real)1 : (real)-1;
            const real nv1 = ((Mir[1][s >> 5] >> (s & 31)) & 1u) ? (real)1 : (real)-1;
            nsum = nv0 + nv1;
            /* the sum if the pending neighbour attempt flips this site */
            nsumAlt = hypNb == 0 ? -nv0 + nv1 : nv0 - nv1;
        }
        auto decideWith = [&](real f, real ns_) {
            real thr;
            if (!a.sa) {
                /* CPUDenseGraphAnnealer.cpp:258-262, same operation order */
                real dE = a.twoDivM * q * (hx + (real)2. * f);
                dE -= q * ns_ * a.coef;
                /* evaluated for every lane and selected (no divergent branch) */
                const real e = Expo<real>::eval(-dE * a.beta, expTab);
                thr = (dE < (real)0.) ? (real)1. : e;
            } else {
                /* :348-349 */
                real dE = (real)2. * q * (hx + (real)2. * f);
                const real e = Expo<real>::eval(-dE * a.invKT, expTab);
                thr = (dE < (real)0.) ? (real)1. : e;
            }
            return thr > u;
        };
        auto decide = [&](real f) { return decideWith(f, nsum); };
        bool accLocal = false;
        real fLane = 0; /* this lane's field as its decision used it */
        if constexpr (PLAIN) {
            blocked = blocked || hypNb >= 0;
        } else {
            /* decided under both outcomes of the pending attempt: if they
             * agree the decision is final, otherwise the lane waits */
            if constexpr (kHelp) {
                if (helpBusy) {
                    SQ_CYC(h0);
                    while (helpLoad(ctl) != kHelpIdle)
                        ;
                    helpBusy = false;
                    SQ_CYC(h1);
                    SQ_CYC_ADD(cyc[11], h0, h1);
                }
            }
            real f = F[s];
            if constexpr (kMultiCap) {
#pragma unroll
                for (int q = 0; q < kMaxPend; ++q)
                    if (q < nPend)
                        f = addScaled(f, pendD[q], pendCorr[q]);
            } else if constexpr (kDefer) {
                if (rowPending)
                    f = addScaled(f, rowD, rowCorr);
            }
            accLocal = decide(f);
            if (hypNb >= 0 && decideWith(f, nsumAlt) != accLocal)
                blocked = true;
            fLane = f;
        }
        (void)fLane;
        (void)blockedMulti;
        const unsigned long long stopMask = __ballot(!valid || blocked);
        const int firstStop = uni(firstLane(stopMask));
        int c = firstStop, flipLane = -1;
        unsigned long long flipMask = 0; /* the committed flips, by lane */
        real dFlip = 0;                  /* a flipped lane's row factor -2 q (q before its flip) */
        int nCand = 0;                   /* lanes that would accept (the next batch's mode) */
        if constexpr (PLAIN) {
            /* reference algorithm: field = J[s] . q_y recomputed per attempt,
             * as coalesced wave-wide dots, eight attempts at a time, stopping
             * at the first accepted flip */
            real f = 0;
            for (int g = 0; g < firstStop && flipLane < 0; g += kDotGroup) {
                real acc[kDotGroup];
                uint32_t site[kDotGroup];
#pragma unroll
                for (int t = 0; t < kDotGroup; ++t) {
                    acc[t] = 0;
                    site[t] = __shfl(s, (g + t) & (kWave - 1));
                }
                const int nDots = firstStop - g < kDotGroup ? firstStop - g : kDotGroup;
#pragma unroll 4
                for (int b = 0; b < Np; b += kWave * 4) { /* uniform bound: a scalar loop */
                    const int j = b + lane * 4;
                    const uint32_t qw = Qb[j >> 5] >> (j & 31);
#pragma unroll
                    for (int t = 0; t < kDotGroup; ++t) {
                        if (t < nDots) {
                            real v[4];
                            loadJ4(a.J + (long long)site[t] * Np + j, v);
                            acc[t] += ((qw & 1u) ? v[0] : -v[0]) + ((qw & 2u) ? v[1] : -v[1]) +
                                      ((qw & 4u) ? v[2] : -v[2]) + ((qw & 8u) ? v[3] : -v[3]);
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < kDotGroup; ++t) {
                    real v = acc[t];
                    for (int off = kWave / 2; off > 0; off >>= 1)
                        v += __shfl_xor(v, off);
                    if (lane == g + t)
                        f = v;
                }
                const bool inGroup = lane >= g && lane < g + nDots;
                const unsigned long long accMask = __ballot(inGroup && decide(f));
                if (accMask)
                    flipLane = uni(firstLane(accMask));
            }
            if (flipLane >= 0) {
                c = flipLane + 1;
                flipMask = 1ull << flipLane;
            }
        } else if (kMultiCap && multi) {
            /* lanes decide in order; after each accepted flip (lane fa, site
             * x, factor d = -2 q) the later lanes' fields get the serial
             * update F[s] = F[s] + d J[x][s] (J[x][s] = sub[fa][lane]), a
             * later attempt at site x sees the flipped spin, and their
             * decisions and two-outcome checks are redone.  Lanes before the
             * next accepted one are confirmed rejections; the batch ends at
             * the first lane that is blocked, past the window or past the
             * submatrix (kMultiLanes). */
            int j = 0;
            c = kWave;
            for (;;) {
                const bool act = lane >= j;
                const int fs = uni(firstLane(__ballot(act && (!valid || blocked || lane >= kMultiLanes))));
                const int fa = uni(firstLane(__ballot(act && valid && !blocked && accLocal)));
                if (fa >= fs) {
                    c = fs;
                    break;
                }
                flipMask |= 1ull << fa;
                const uint32_t xa = (uint32_t)__builtin_amdgcn_readlane((int)s, fa);
                const real da = __builtin_amdgcn_readlane((int)(q > (real)0), fa) ? (real)-2. : (real)2.;
                if (lane == fa)
                    dFlip = da;
                if (lane > fa) {
                    fLane = addScaled(fLane, da, sub[fa * kWave + lane]);
                    if (s == xa)
                        q = -q;
                    accLocal = decide(fLane);
                    blocked = blockedMulti || (hypNb >= 0 && decideWith(fLane, nsumAlt) != accLocal);
                }
                j = fa + 1;
                if (j >= kWave)
                    break;
            }
            flipLane = flipMask ? uni(63 - __builtin_clzll(flipMask)) : -1; /* the last flip */
            nCand = __popcll(flipMask);
        } else {
            const bool acc = accLocal;
            const unsigned long long accMask = __ballot(valid && !blocked && acc);
            const int firstAcc = uni(firstLane(accMask));
            nCand = __popcll(accMask);
            if (firstAcc < firstStop) {
                c = firstAcc + 1;
                flipLane = firstAcc;
                flipMask = 1ull << firstAcc;
                if (lane == firstAcc)
                    dFlip = qpos ? (real)-2. : (real)2.;
            }
        }
        waveSync();
        SQ_CYC(c3);
        SQ_CYC_ADD(cyc[2], c2, c3);

        /* clear the pending-site table */
        if (!a.sa) {
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                /* every lane clears its slot (no exec masking): every entry of
                 * this batch is cleared, and a slot nobody entered is empty */
                tbl[pendSite[nb][0] & tblMask] = kPendEmpty;
                tbl[pendSite[nb][1] & tblMask] = kPendEmpty;
            }
        }

        if (c == 0) {
            /* lane 0 waits on a neighbour's serially earlier attempt: poll the
             * earliest pending entry at lane 0's site (or, when the window does
             * not reach lane 0's attempt, the first undecided entry) */
            ++nBlocked;
            if (!a.sa) {
                const uint32_t s0 = __shfl(s, 0);
                int waitNb = -1, waitE = 0x7fffffff;
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) {
                    const int base = curBase[nb];
                    const int Kk = k + delta[nb];
                    const int e0 = base + lane, e1 = base + kWave + lane;
                    const bool m0 = pendSetNow[nb][0] && pendSite[nb][0] == s0 && e0 < Kk;
                    const bool m1 = pendSetNow[nb][1] && pendSite[nb][1] == s0 && e1 < Kk;
                    const unsigned long long b0 = __ballot(m0), b1 = __ballot(m1);
                    int e = b0 ? base + firstLane(b0) : (b1 ? base + kWave + firstLane(b1) : 0x7fffffff);
                    if (e == 0x7fffffff && limit <= 0 && applied[nb] < Kk)
                        e = applied[nb]; /* window short of Kk: the first undecided entry */
                    if (e < waitE) {
                        waitE = e;
                        waitNb = nb;
                    }
                }
                if (waitNb >= 0 && waitE < N) {
                    const uint32_t *dp = a.D + (long long)nbr[waitNb] * N + waitE;
                    while ((loadShared32(dp) >> 1) != tag) {
                        if (++spins > a.spinLimit)
                            break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
            }
            if (++spins > a.spinLimit) {
                atomicExch(a.abortFlag, 1);
                if (a.abortHost) /* a vector store, system scope: the host reads it after the step */
                    __hip_atomic_store(a.abortHost, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                aborted = true;
                break;
            }
            if ((spins & 63) == 0 && __hip_atomic_load(a.abortFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                aborted = true;
                break;
            }
            prefetch(true);
            waveSync();
            {
                SQ_CYC(cb);
                SQ_CYC_ADD(cyc[5], c0, cb);
            }
            continue;
        }
        spins = 0;
        ++nBatches;

        /* publish the committed decisions */
        if (lane < c) /* lane 0 always stores: one op */
            storeShared32(a.D + yN + k + lane, (tag << 1) | (uint32_t)((flipMask >> lane) & 1ull));
        afterGroup += 1;

        SQ_CYC(c4);
        if constexpr (kMultiCap) {
            if (nPend > 0)
                applyPendingRows();
        } else if constexpr (kHelp) {
            if (rowPending) {
                /* the row landed before this batch's evaluation (vmcnt) */
                helpStore(ctl, helpApply(curStage, rowD < (real)0.));
                helpBusy = true;
                rowPending = false;
            }
        } else if constexpr (kDefer) {
            if (rowPending)
                applyPendingRow();
        }
        SQ_CYC(r1);
        SQ_CYC_ADD(cyc[9], c4, r1);
        if (flipLane >= 0) {
            const uint32_t x = __shfl(s, flipLane);
            const bool qx = __shfl((int)qpos, flipLane) != 0;
            const int nFl = kMultiCap ? uni(__popcll(flipMask)) : 1;
            nFlips += nFl;
            /* every flipped lane toggles its own site's bit (a site flipped
             * twice in one batch toggles twice) */
            if ((flipMask >> lane) & 1ull)
                atomicXor(&Qb[s >> 5], 1u << (s & 31));
            (void)x;
            (void)qx;
            if constexpr (FP8) {
                /* the same deferred update from the e4m3 row (ld8 bytes, a
                 * multiple of 1 KiB) and its scale, rowD = -+2 * scale */
                rowSign = qx ? -2.f : 2.f;
                const uint8_t *row = a.J8 + (long long)x * a.ld8;
                waveSync();
                ldsWait(); /* this wave's reads of `stage` and the scale are done */
                dmaRunx4(row + lane * 16, stageL, a.ld8 / (kWave * 16));
                if (lane == 0)
                    dmaDword<false>(a.rowScale + x, scaleL);
                afterGroup += rowOps;
                rowPending = true;
            } else if constexpr (kMultiCap) {
                /* F[y][:] += -2 q_x J[x][:] (J symmetric) for the flips in lane
                 * order.  The last min(nFl, nBufs) rows are deferred: they
                 * stream into stage buffers 0, 1, ..., the next batch's lanes
                 * add them to their fields in that order, and the next commit
                 * applies them in one pass.  Earlier rows (batches with more
                 * flips than buffers) stream through the buffers and are
                 * applied here one by one: row r is in buffer r % nBufs, a
                 * buffer is refilled only after its row was applied, and
                 * vmcnt retires in issue order, so row r has landed once at
                 * most (issued - 1 - r) rows' ops are outstanding ((nBufs - 1)
                 * rows' ops fit the counter). */
                waveSync();
                ldsWait(); /* this wave's reads of the stage buffers are done */
                const int nDefer = nFl < nBufs ? nFl : nBufs;
                const int nImm = nFl - nDefer;
                unsigned long long issueMask = flipMask, applyMask = flipMask;
                /* buffer indices wrap by compare (r % nBufs without a divide) */
                int issued = 0, applied = 0, issueBuf = 0, applyBuf = 0;
                auto issueNext = [&]() {
                    const int fl = uni(firstLane(issueMask));
                    issueMask &= issueMask - 1;
                    const uint32_t xr = (uint32_t)__builtin_amdgcn_readlane((int)s, fl);
                    const real *row = a.J + (long long)xr * Np;
                    dmaRunx4(reinterpret_cast<const unsigned char *>(row + lane * 4), stageBufL(issueBuf), rowOps);
                    afterGroup += rowOps;
                    ++issued;
                    issueBuf = (issueBuf + 1 == nBufs) ? 0 : issueBuf + 1;
                };
                while (issued < nImm && issued < nBufs)
                    issueNext();
                /* rows land in issue order; two landed rows are added in one
                 * pass over F (each element takes them in flip order), while
                 * the next ones stream in */
                while (applied < nImm) {
                    if (applied + 1 < nImm) {
                        /* rows applied and applied + 1 (issued >= applied + 2:
                         * nBufs >= 2) */
                        SQ_CYC(w0);
                        vmWait((issued - 2 - applied) * rowOps);
                        SQ_CYC(w1);
                        SQ_CYC_ADD(cyc[12], w0, w1);
                        const int fl0 = uni(firstLane(applyMask));
                        applyMask &= applyMask - 1;
                        const int fl1 = uni(firstLane(applyMask));
                        applyMask &= applyMask - 1;
                        const int b1 = (applyBuf + 1 == nBufs) ? 0 : applyBuf + 1;
                        const real *const sbs[2] = {stageBuf(applyBuf), stageBuf(b1)};
                        const real ds[2] = {readLaneF(dFlip, fl0), readLaneF(dFlip, fl1)};
                        addStagedRows<real, 2>(F, sbs, ds, Np, lane);
                        applied += 2;
                        applyBuf = (b1 + 1 == nBufs) ? 0 : b1 + 1;
                    } else {
                        SQ_CYC(w0);
                        vmWait((issued - 1 - applied) * rowOps);
                        SQ_CYC(w1);
                        SQ_CYC_ADD(cyc[12], w0, w1);
                        const int fl = uni(firstLane(applyMask));
                        applyMask &= applyMask - 1;
                        applyRow(applyBuf, readLaneF(dFlip, fl));
                        ++applied;
                        applyBuf = (applyBuf + 1 == nBufs) ? 0 : applyBuf + 1;
                    }
                    if (issued < nImm) {
                        ldsWait(); /* the applied buffers' reads are done */
                        while (issued < nImm && issued < applied + nBufs)
                            issueNext();
                    }
                }
                ldsWait(); /* the applied buffers' reads are done */
                issueBuf = 0;
#pragma unroll
                for (int q = 0; q < kMaxPend; ++q) {
                    if (q < nDefer) {
                        const int fl = uni(firstLane(applyMask));
                        applyMask &= applyMask - 1;
                        pendD[q] = readLaneF(dFlip, fl);
                        issueNext();
                    }
                }
                nPend = nDefer;
            } else if constexpr (kDefer) {
                /* F[y][:] += -2 q_x J[x][:] (J symmetric), deferred: the row
                 * streams into `stage`, 1 KiB (16 B per lane) per op */
                rowD = qx ? (real)-2. : (real)2.;
                const real *row = a.J + (long long)x * Np;
                waveSync();
                ldsWait(); /* this wave's reads of `stage` are done */
                uint32_t dst = stageL;
                if constexpr (kHelp) {
                    /* the other buffer than the one the helper may be reading */
                    curStage ^= 1;
                    dst = stageBufL(curStage);
                }
                dmaRunx4(reinterpret_cast<const unsigned char *>(row + lane * 4), dst, rowOps);
                afterGroup += rowOps;
                rowPending = true;
            } else if constexpr (!PLAIN) {
                /* F[y][:] += -2 q_x J[:][x] = -2 q_x J[x][:]  (J symmetric);
                 * the row's loads go out together, then the LDS updates */
                const real d = qx ? (real)-2. : (real)2.;
                const real *row = a.J + (long long)x * Np;
                constexpr int kRowChunk = sizeof(real) == 4 ? 16 : 8;
                for (int j0 = lane * 4; j0 < Np; j0 += kWave * 4 * kRowChunk) {
                    real v[kRowChunk][4];
#pragma unroll
                    for (int c4 = 0; c4 < kRowChunk; ++c4) {
                        const int j = j0 + c4 * kWave * 4;
                        if (j < Np)
                            loadJ4(row + j, v[c4]);
                    }
#pragma unroll
                    for (int c4 = 0; c4 < kRowChunk; ++c4) {
                        const int j = j0 + c4 * kWave * 4;
                        if (j < Np) {
                            real fv[4];
                            ldsLoad4(F + j, fv);
#pragma unroll
                            for (int t = 0; t < 4; ++t)
                                fv[t] = addScaled(fv[t], d, v[c4][t]);
                            ldsStore4(F + j, fv);
                        }
                    }
                }
            }
        }
        SQ_CYC(r2);
        SQ_CYC_ADD(cyc[10], r1, r2);
        if constexpr (kMultiCap) {
            /* enough lanes accepting (flips of a multi batch, or candidates
             * of a single-flip batch): the next batch resolves its flips in
             * order instead of ending at the first */
            multi = nBufs >= 2 && nCand >= uni(a.multiMinCand);
        }
        k = uni(k + c);
        waveSync();
        SQ_CYC(c5);
        SQ_CYC_ADD(cyc[3], c3, c5);
        SQ_CYC_ADD(cyc[4], c4, c5);
    }

    /* nothing may land in this region after the wave leaves it */
    vmWaitImm<0>();
    if constexpr (kHelp) {
        if (helpBusy)
            while (helpLoad(ctl) != kHelpIdle)
                ;
    }
    if constexpr (kMultiCap) {
        if (nPend > 0 && !aborted) {
            applyPendingRows();
            waveSync();
        }
    } else if constexpr (kDefer) {
        if (rowPending && !aborted) {
            if constexpr (FP8)
                rowD = rowSign * scaleSlot[0];
            applyPendingRow();
            waveSync();
        }
    }
    if (!aborted) {
        if (!PLAIN) {
#pragma unroll 4
            for (int b = 0; b < Np; b += kWave * 4) {
                real v[4];
                ldsLoad4(F + b + lane * 4, v);
                ldsStore4(a.F + (long long)y * Np + b + lane * 4, v); /* a plain vector store, here to global */
            }
        }
        for (int b = 0; b < Nw; b += kWave)
            if (b + lane < Nw)
                a.Qbits[(long long)y * Nw + b + lane] = Qb[b + lane];
    }
    if constexpr (kHelp)
        helpStore(ctl, kHelpExit);
    if (lane == 0 && a.counters) {
        atomicAdd(&a.counters[0], (unsigned long long)nFlips);
        atomicAdd(&a.counters[1], (unsigned long long)nBatches);
        atomicAdd(&a.counters[2], (unsigned long long)nBlocked);
#ifdef SQAOD_SWEEP_CYCLES
        for (int t = 0; t < 13; ++t)
            atomicAdd(&a.counters[3 + t], cyc[t]);
#endif
    }
}

template <class real, int MODE, bool MULTI = false, bool HELPER = false>
__global__ __launch_bounds__(kMaxSlicesPerBlock *kWave) void denseSweepKernel(SweepArgs<real> a, int slicesPerBlock,
                                                                                int xcdRemap) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int w = uni(threadIdx.x / kWave); /* the wave index, uniform per wave */

    /* XCD-aware placement: consecutive slices on the same XCD (blocks are
     * dispatched round-robin over the 8 XCDs). */
    int b = blockIdx.x;
    if (xcdRemap) {
        int per = gridDim.x / 8;
        b = (blockIdx.x % 8) * per + blockIdx.x / 8;
    }
    /* HELPER: one slice per block, wave 0 sweeps it and wave 1 applies rows */
    const int y = HELPER ? b : b * slicesPerBlock + w;

    /* LDS carve-up: exp table, h, then one region per slice */
    uint64_t *expTab = reinterpret_cast<uint64_t *>(lds);
    real *hS = reinterpret_cast<real *>(lds + kExpTabBytes);
    const size_t waveBytes = sweepWaveBytes<real, MODE>(a.Np, a.ld8, MULTI ? a.multiBufs : 0, HELPER);
    unsigned char *p = lds + kExpTabBytes + (size_t)a.Np * sizeof(real) + waveBytes * (HELPER ? 0 : w);
    if (threadIdx.x < 32)
        expTab[threadIdx.x] = sq_exp2_tab32[threadIdx.x];
    if constexpr (HELPER) {
        if (threadIdx.x == 0)
            *helperCtl(a, p) = kHelpIdle;
    }
    {
        /* h by DMA, the block's waves taking 1 KiB pieces in turn (h holds Np
         * reals, a multiple of 1 KiB) */
        const int lane = threadIdx.x & (kWave - 1), nW = uni(blockDim.x / kWave);
        const unsigned char *src = reinterpret_cast<const unsigned char *>(a.h);
        const uint32_t hL = ldsAddress(hS);
        for (int j0 = w * kWave * 16; j0 < a.Np * (int)sizeof(real); j0 += nW * kWave * 16)
            dmaDwordx4(src + j0 + lane * 16, hL + j0);
        vmWaitImm<0>();
    }
    __syncthreads();
    if (y >= a.m)
        return;
    if constexpr (HELPER) {
        if (w == 1) {
            helperLoop<real>(a, p);
            return;
        }
    }
    sweepSlice<real, MODE, MULTI, HELPER>(a, y, expTab, hS, p);
}

/* Many instances in one persistent launch: every wave takes (instance,
 * slice) items from a queue in item order until it is empty, loads the
 * instance's h into its own LDS region and runs the slice's sweep.  Items of
 * one instance are consecutive, so the waves running them start together;
 * a wave that waits on a neighbour slice waits only on items already taken
 * (by running waves) or next in the queue.  With at least max(m) waves
 * resident every instance therefore completes: the items taken are a prefix
 * of the queue, every instance before the last started one has all its
 * slices running, and the last one's remaining slices are taken as soon as
 * any earlier instance frees its waves.  regionBytes is the per-wave region
 * (h + sweepWaveBytes) for the largest instance. */
template <class real, int MODE>
__global__ __launch_bounds__(kMaxSlicesPerBlock *kWave) void denseSweepBatchKernel(const SweepArgs<real> *args,
                                                                                     const uint2 *items, int nItems,
                                                                                     int *queue, int regionBytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = uni(threadIdx.x / kWave); /* the wave index, uniform per wave */
    uint64_t *expTab = reinterpret_cast<uint64_t *>(lds);
    unsigned char *region = lds + kExpTabBytes + (size_t)regionBytes * w;
    if (threadIdx.x < 32)
        expTab[threadIdx.x] = sq_exp2_tab32[threadIdx.x];
    __syncthreads();
    for (;;) {
        int item = 0;
        if (lane == 0)
            item = atomicAdd(queue, 1);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0));
        if (item >= nItems)
            break;
        const uint2 it = items[item];
        const SweepArgs<real> a = args[__builtin_amdgcn_readfirstlane((int)it.x)];
        real *hS = reinterpret_cast<real *>(region);
        {
            const unsigned char *src = reinterpret_cast<const unsigned char *>(a.h);
            const uint32_t hL = ldsAddress(hS);
            for (int j0 = 0; j0 < a.Np * (int)sizeof(real); j0 += kWave * 16)
                dmaDwordx4(src + j0 + lane * 16, hL + j0);
            vmWaitImm<0>();
        }
        sweepSlice<real, MODE>(a, uni((int)it.y), expTab, hS, region + (size_t)a.Np * sizeof(real));
        waveSync();
    }
}

__global__ void resetQueueKernel(int *queue) { *queue = 0; }

__global__ void fp8DecodeKernel(float *out) {
    const int t = threadIdx.x; /* 64 threads, 4 codes each */
    float v[4];
    fp8x4ToFloat((uint32_t)(4 * t) | ((uint32_t)(4 * t + 1) << 8) | ((uint32_t)(4 * t + 2) << 16) |
                     ((uint32_t)(4 * t + 3) << 24),
                 v);
    for (int i = 0; i < 4; ++i)
        out[4 * t + i] = v[i];
}

}  // namespace

void fp8DecodeTable(float out[256]) {
    float *d = nullptr;
    HIP_CHECK(hipMalloc((void **)&d, 256 * sizeof(float)));
    fp8DecodeKernel<<<1, 64>>>(d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpy(out, d, 256 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    HIP_CHECK(e);
}

namespace {
size_t waveBytesFor(int Np, size_t realSize, bool plain, int fieldFormat, int ld8, int multiBufs = 0,
                    bool helper = false) {
    if (realSize == 8)
        return plain ? sweepWaveBytes<double, kModePlain>(Np, ld8) : sweepWaveBytes<double, kModeLocal>(Np, ld8);
    if (plain)
        return sweepWaveBytes<float, kModePlain>(Np, ld8);
    return fieldFormat == fieldFP8 ? sweepWaveBytes<float, kModeFP8>(Np, ld8)
                                   : sweepWaveBytes<float, kModeLocal>(Np, ld8, multiBufs, helper);
}

template <class real>
void *batchKernelFor(bool plain, int fieldFormat) {
    if (plain)
        return (void *)denseSweepBatchKernel<real, kModePlain>;
    if constexpr (sizeof(real) == 4)
        if (fieldFormat == fieldFP8)
            return (void *)denseSweepBatchKernel<real, kModeFP8>;
    return (void *)denseSweepBatchKernel<real, kModeLocal>;
}
}  // namespace

size_t denseSweepLdsBytes(int Np, int slicesPerBlock, size_t realSize, bool plain, int fieldFormat, int ld8,
                          int multiBufs, bool helper) {
    return kExpTabBytes + (size_t)Np * realSize +
           waveBytesFor(Np, realSize, plain, fieldFormat, ld8, multiBufs, helper) * slicesPerBlock;
}

bool denseSweepHelperFits(int Np, int slicesPerBlock, size_t realSize, bool plain, int fieldFormat,
                          size_t maxLds) {
    return realSize == 4 && !plain && fieldFormat != fieldFP8 && slicesPerBlock == 1 &&
           denseSweepLdsBytes(Np, 1, realSize, plain, fieldFormat, 0, 0, true) <= maxLds;
}

int denseSweepMultiBufs(int Np, int slicesPerBlock, size_t realSize, bool plain, int fieldFormat, size_t maxLds) {
    if (realSize != 4 || plain || fieldFormat == fieldFP8)
        return 0;
    const int rowOps = Np * 4 / (kWave * 16); /* 1 KiB DMA ops per row */
    for (int b = 4; b >= 2; --b)
        if ((b - 1) * rowOps <= 63 &&
            denseSweepLdsBytes(Np, slicesPerBlock, realSize, plain, fieldFormat, 0, b) <= maxLds)
            return b;
    return 0;
}

size_t denseSweepBatchRegionBytes(int maxNp, int maxLd8, size_t realSize, bool plain, int fieldFormat) {
    return (size_t)maxNp * realSize + waveBytesFor(maxNp, realSize, plain, fieldFormat, maxLd8);
}

template <class real>
int denseSweepBatchBlocks(size_t regionBytes, int wavesPerBlock, bool plain, int fieldFormat, int numCUs) {
    SQ_THROW_IF(wavesPerBlock < 1 || wavesPerBlock > kMaxSlicesPerBlock, "waves per block out of range (1..%d)",
                kMaxSlicesPerBlock);
    const size_t lds = kExpTabBytes + regionBytes * wavesPerBlock;
    void *fn = batchKernelFor<real>(plain, fieldFormat);
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int perCU = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, wavesPerBlock * kWave, lds));
    return perCU * numCUs;
}

template <class real>
void denseSweepBatch(const SweepArgs<real> *d_args, const uint2 *d_items, int nItems, int *d_queue, bool plain,
                     int fieldFormat, size_t regionBytes, int wavesPerBlock, int nBlocks, hipStream_t stream) {
    SQ_THROW_IF(nItems <= 0 || nBlocks <= 0, "batch sweep: nothing to run");
    SQ_THROW_IF(regionBytes % 16 != 0, "batch sweep: region of %zu bytes is not 16-byte aligned", regionBytes);
    /* the LDS attribute was set by denseSweepBatchBlocks (nothing but stream
     * work here, so the launch can be captured into a graph) */
    const size_t lds = kExpTabBytes + regionBytes * wavesPerBlock;
    resetQueueKernel<<<1, 1, 0, stream>>>(d_queue);
    HIP_CHECK(hipGetLastError());
    const int rb = (int)regionBytes;
    if (plain)
        denseSweepBatchKernel<real, kModePlain><<<nBlocks, wavesPerBlock * kWave, lds, stream>>>(d_args, d_items, nItems,
                                                                                                d_queue, rb);
    else if (sizeof(real) == 4 && fieldFormat == fieldFP8)
        denseSweepBatchKernel<real, (sizeof(real) == 4 ? kModeFP8 : kModeLocal)>
            <<<nBlocks, wavesPerBlock * kWave, lds, stream>>>(d_args, d_items, nItems, d_queue, rb);
    else
        denseSweepBatchKernel<real, kModeLocal><<<nBlocks, wavesPerBlock * kWave, lds, stream>>>(d_args, d_items, nItems,
                                                                                                d_queue, rb);
    HIP_CHECK(hipGetLastError());
}

template int denseSweepBatchBlocks<float>(size_t, int, bool, int, int);
template int denseSweepBatchBlocks<double>(size_t, int, bool, int, int);
template void denseSweepBatch<float>(const SweepArgs<float> *, const uint2 *, int, int *, bool, int, size_t, int, int,
                                     hipStream_t);
template void denseSweepBatch<double>(const SweepArgs<double> *, const uint2 *, int, int *, bool, int, size_t, int,
                                      int, hipStream_t);

/* Whether every block of a launch of fn fits on the device at once. */
static bool fitsAtOnce(const void *fn, int threads, size_t lds, int nBlocks, int numCUs) {
    static std::mutex mu;
    static std::map<std::tuple<const void *, int, size_t>, int> perCU;
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_tuple(fn, threads, lds);
    auto it = perCU.find(key);
    if (it == perCU.end()) {
        int n = 0;
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, threads, lds));
        it = perCU.emplace(key, n).first;
    }
    return (long long)it->second * numCUs >= nBlocks;
}

/* The sweep's waves spin on each other, so all of them must be resident.
 * A cooperative launch guarantees that, at the price of the runtime's queue
 * synchronisation around it (about 20 us per launch here, 5% of a step at
 * N=4096, m=256).  A direct launch is used instead when allowDirect (the
 * caller's stream is the only one on this GPU launching sweeps in this
 * process) and every block fits on the device at once: other kernels that
 * hold CUs finish, so the dispatched blocks all become resident.  If they
 * cannot (another process's spinning kernels), the spin limit turns the wait
 * into an error. */
template <class real>
void denseSweep(const SweepArgs<real> &args, bool plain, int fieldFormat, int slicesPerBlock, bool allowDirect,
                int numCUs, hipStream_t stream) {
    SQ_THROW_IF(args.Np % 256 != 0, "padded N must be a multiple of 256");
    SQ_THROW_IF(slicesPerBlock < 1 || slicesPerBlock > kMaxSlicesPerBlock, "slices per block out of range (1..%d)",
                kMaxSlicesPerBlock);
    const bool fp8 = !plain && fieldFormat == fieldFP8;
    SQ_THROW_IF(fp8 && (sizeof(real) != 4 || !args.J8 || !args.rowScale || args.ld8 < args.Np || args.ld8 % 1024),
                "fp8 sweep: needs a float solver and e4m3 rows with a 1 KiB-multiple stride");
    /* the helper-wave kernel: float local-field form, single-flip batches,
     * one slice per block */
    const bool helper =
        args.helperWave && sizeof(real) == 4 && !plain && !fp8 && args.multiBufs < 2 && slicesPerBlock == 1;
    SQ_THROW_IF(args.helperWave && !helper && sizeof(real) == 4 && !plain && !fp8 && args.multiBufs < 2,
                "helper-wave sweep: needs one slice per block");
    int nBlocks = divru(args.m, slicesPerBlock);
    size_t lds = denseSweepLdsBytes(args.Np, slicesPerBlock, sizeof(real), plain, fieldFormat, args.ld8,
                                    args.multiBufs, helper);
    int xcdRemap = (nBlocks % 8 == 0 && nBlocks >= 16) ? 1 : 0;

    void *fn = plain ? (void *)denseSweepKernel<real, kModePlain> : (void *)denseSweepKernel<real, kModeLocal>;
    if constexpr (sizeof(real) == 4) {
        if (fp8)
            fn = (void *)denseSweepKernel<real, kModeFP8>;
        else if (!plain && args.multiBufs >= 2)
            fn = (void *)denseSweepKernel<real, kModeLocal, true>;
        else if (helper)
            fn = (void *)denseSweepKernel<real, kModeLocal, false, true>;
    }
    const int threads = helper ? 2 * kWave : slicesPerBlock * kWave;
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    SweepArgs<real> a = args;
    void *kargs[] = {(void *)&a, (void *)&slicesPerBlock, (void *)&xcdRemap};
    if (allowDirect && fitsAtOnce(fn, threads, lds, nBlocks, numCUs))
        HIP_CHECK(hipLaunchKernel(fn, dim3(nBlocks), dim3(threads), kargs, (unsigned)lds, stream));
    else
        HIP_CHECK(hipLaunchCooperativeKernel(fn, dim3(nBlocks), dim3(threads), kargs, (unsigned)lds, stream));
}

template void denseSweep<float>(const SweepArgs<float> &, bool, int, int, bool, int, hipStream_t);
template void denseSweep<double>(const SweepArgs<double> &, bool, int, int, bool, int, hipStream_t);

}  // namespace sqaod_rocm
