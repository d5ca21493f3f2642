/*
 * Host-side launchers of the sqaod-rocm HIP kernels (implemented in the .hip
 * sources of this directory).  All launchers are asynchronous on `stream`.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sqaod_rocm {

/* ---------------- random streams (rng.hip) ---------------- */

/* MT19937 state resident in device memory: 624 words and the index of the
 * next word to temper (624 = a twist is due). */
struct MTState {
    uint32_t mt[624];
    uint32_t idx;
};

/* Host-side init_genrand(seed & 0xffffffff) (Random.cpp:70-83). */
void mtSeedHost(MTState *st, unsigned long long seed);

/* Appends n tempered MT19937 outputs to out[0..n), continuing the device
 * state; one workgroup, order identical to the serial host generator. */
void mtGenerate(MTState *d_state, uint32_t *d_out, size_t n, hipStream_t stream);

/* Ordering of attempts inside one sweep, i.e. which stream position feeds
 * attempt (slice y, iteration k). */
enum AttemptOrder {
    orderColoring = 0, /* per iteration k: even slices ascending, then odd (CPUDenseGraphAnnealer.cpp:282-292) */
    orderSliceMajor = 1, /* slice by slice, N attempts each (SA naive, :366-369) */
};

/* Serial-stream mapping: attempt (y, k) consumes words [a*w, a*w + w) with
 * a its position in `order` and w = 2 (float: site, u) or 3 (double: site,
 * 2 draws for u).  Writes sites X[y*N+k] = word % N and uniforms U[y*N+k]. */
template <class real>
void streamToAttempts(const uint32_t *d_words, int N, int m, int order, uint32_t *d_X, real *d_U,
                      hipStream_t stream);

/* sqaod.cpu's OpenMP coloring step with T workers
 * (CPUDenseGraphAnnealer.cpp:303-339): worker w draws from its own generator
 * d_states[w] (seeded s + 17 w, :29-33) for the slices of each colour in its
 * static chunk, worker 0 also for slice m-1 of odd m between the colours.
 * Generates every worker's words of one step (worker w at
 * d_words + w * mtWorkerStride) and writes X, U in (slice, attempt) order. */
unsigned long long mtWorkerStride(int N, int m, int T, int perAttempt);
template <class real>
void workerStreamsToAttempts(MTState *d_states, uint32_t *d_words, int N, int m, int T, uint32_t *d_X, real *d_U,
                             hipStream_t stream);

/* Counter-based stream (Philox4x32-10): attempt (y, k) of sweep `sweepIdx`
 * gets one Philox block keyed by the seed. */
template <class real>
void philoxAttempts(unsigned long long seed, unsigned long long sweepIdx, int N, int m, uint32_t *d_X,
                    real *d_U, hipStream_t stream);

/* The same streams for many sweeps in one launch (values identical to
 * philoxAttempts with the job's seed and sweep index): job j owns the
 * kAttemptBlock-thread blocks [blockStart[j], blockStart[j] + ceil(N m / kAttemptBlock)). */
constexpr int kAttemptBlock = 256;
struct AttemptStreamJob {
    unsigned long long seed, sweepIdx;
    int N, m;
    uint32_t *X;
    void *U; /* real */
};
template <class real>
void philoxAttemptsGrouped(const AttemptStreamJob *d_jobs, const int *d_blockStart, int nJobs, int totalBlocks,
                           hipStream_t stream);

/* randomizeSpin from a word stream: bit (y, x) = word[y*rowStride+x] & 1
 * (CPUDenseGraphAnnealer.cpp:152-161: randInt(2) ? +1 : -1; the bipartite
 * annealer interleaves the q0 and q1 rows of each slice,
 * CPUBipartiteGraphAnnealer.cpp:183-192).  Qbits is m x Nw words, bit x%32 of
 * word x/32; padding bits are cleared. */
void spinsFromWords(const uint32_t *d_words, long long rowStride, int N, int m, int Nw, uint32_t *d_Qbits,
                    hipStream_t stream);
/* randomizeSpin from Philox (domain separated from the attempt streams). */
void spinsFromPhilox(unsigned long long seed, unsigned long long salt, int N, int m, int Nw, uint32_t *d_Qbits,
                     hipStream_t stream);

/* int8 spins (m x N, +-1) <-> packed bits */
void packSpins(const int8_t *d_q, int N, int m, int Nw, uint32_t *d_Qbits, hipStream_t stream);
void unpackSpins(const uint32_t *d_Qbits, int N, int m, int Nw, int8_t *d_q, hipStream_t stream);

/* ---------------- dense local fields (dense_field.hip) ---------------- */

enum FieldFormat {
    fieldF32 = 0,    /* J stored as real */
    fieldBF16x2 = 1, /* J ~ hi + lo, both bf16, hi = bf16(J), lo = bf16(J - hi) */
    fieldFP8 = 2     /* J ~ J' = 2^a_i e4m3 per row i (quantizeFP8) */
};

/* Jsplit = float(bf16(J)) + float(bf16(J - bf16(J))) elementwise (exact in
 * fp32): the J the bf16x2 sweep's row updates use, so that incremental fields
 * stay consistent with the split GEMM. */
void splitBF16(const float *d_J, size_t count, float *d_Jsplit, hipStream_t stream);

/* FP8 form of a symmetric J (N x ld floats): each row i gets a power-of-two
 * scale 2^a_i (the smallest with max_j |J[i][j]| <= 448 2^a_i, raised to
 * within 15 binades of the largest), and J'[i][j] is J[i][j] rounded to
 * nearest-even on the e4m3 grid of 2^max(a_i, a_j): J' is symmetric, and
 * J'[i][j] / 2^a_i is an exact e4m3 value in both rows.  Outputs: d_rowExp
 * (N ints), d_scale (ld floats, 1 past N), d_J8 (N x ld8 bytes, e4m3
 * J'/2^a_i, zero past N) and d_Jq = J' (N x ld floats). */
void quantizeFP8(const float *d_J, int N, int ld, int ld8, int *d_rowExp, float *d_scale, uint8_t *d_J8, float *d_Jq,
                 hipStream_t stream);

/* F[y][i] = sum_{j<K} q[y][j] * B[j][i] for y < m, i < nCols (B row-major with
 * row stride ldB >= nCols, F row stride ldF), spins from packed bits (Nw words
 * per row), vector ALU; nCols a multiple of 256.  real = float | double.
 * Dense annealer: B = J (symmetric, N x Np), K = N, nCols = ldB = ldF = Np. */
template <class real>
void fieldRefresh(const real *d_B, int K, int nCols, int ldB, const uint32_t *d_Qbits, int m, int Nw, int ldF,
                  real *d_F, hipStream_t stream);

/* MFMA forms (dense_field_mfma.hip): F[y][c] = sum_k q[y][k] B[c][k] for
 * c < nCols (a multiple of 16), k < K (a multiple of 256, Nw = K/32 spin words
 * per row), fp32 accumulation, F row stride ldF.  B (rows x ld floats, zero
 * past its valid k) is first copied into MFMA fragment order:
 * fieldFragElemsBF16x2 = 2 nCols K uint16 (hi and lo planes) or
 * fieldFragElemsF32 = 3 nCols K uint16 (the fp32 form: B split exactly into
 * three bf16 planes, all on the bf16 MFMA).  The launchers check the buffer sizes
 * they are given.  Dense annealer: B = J, rows = N, ld = nCols = K = Np. */
size_t fieldFragElemsBF16x2(int nCols, int K);
size_t fieldFragElemsF32(int nCols, int K);
void buildFragBF16x2(const float *d_B, int rows, int ld, int nCols, int K, uint16_t *d_frag, size_t fragElems,
                     hipStream_t stream);
void buildFragF32(const float *d_B, int rows, int ld, int nCols, int K, uint16_t *d_frag, size_t fragElems,
                  hipStream_t stream);
void fieldRefreshBF16x2(const uint16_t *d_frag, size_t fragElems, const uint32_t *d_Qbits, int nCols, int K, int m,
                        int Nw, int ldF, float *d_F, hipStream_t stream);
void fieldRefreshF32MFMA(const uint16_t *d_frag, size_t fragElems, const uint32_t *d_Qbits, int nCols, int K,
                         int m, int Nw, int ldF, float *d_F, hipStream_t stream);
/* FP8 form: B8 holds e4m3 bytes (rows x ld8) whose row c is scaled by
 * colScale[c] (a power of two); F[y][c] = colScale[c] sum_k q[y][k] B8[c][k]
 * on v_mfma_scale_f32_16x16x128_f8f6f4 (unit block scales), fp32
 * accumulation.  fieldFragElemsFP8 = nCols K bytes. */
size_t fieldFragElemsFP8(int nCols, int K);
/* One GEMM of a grouped launch (fragments already built; colScale for fp8). */
struct FieldGemmJob {
    const void *frag;
    const float *colScale;
    const uint32_t *Qbits;
    float *F;
    int nCols, K, m, Nw, ldF;
};
/* All jobs (one field format) in one launch: job j owns blocks
 * [blockStart[j], blockStart[j] + fieldGemmJobBlocks(job j)); totalBlocks is the sum.
 * checkFieldGemmJob validates a job's shape on the host. */
void checkFieldGemmJob(const FieldGemmJob &job, int fieldFormat);
/* workgroups of a job in a grouped launch (the bf16 forms give a short
 * k-range four column blocks per workgroup) */
int fieldGemmJobBlocks(const FieldGemmJob &job, int fieldFormat);
void fieldRefreshGrouped(int fieldFormat, const FieldGemmJob *d_jobs, const int *d_blockStart, int nJobs,
                         int totalBlocks, hipStream_t stream);
void buildFragFP8(const uint8_t *d_B8, int rows, int ld8, int nCols, int K, uint8_t *d_frag, size_t fragElems,
                  hipStream_t stream);
void fieldRefreshFP8(const uint8_t *d_frag, size_t fragElems, const float *d_colScale, const uint32_t *d_Qbits,
                     int nCols, int K, int m, int Nw, int ldF, float *d_F, hipStream_t stream);

/* ---------------- bipartite annealer flips (bipartite.hip) ---------------- */

/* One flip phase of a bipartite half-step: slices slice0, slice0 + sliceStep,
 * ... of the annealed part (N spins per slice in Nw words, Nw even), with the
 * half-step's dEmat (m x ldE).  Coloring (sa = 0): dE = twoDivM q (h + dEmat)
 * - q (q_prev + q_next) coef, uniforms at serial position rank(im) * N + iq
 * with rank = im / 2 for even im, nEven + im / 2 for odd.  SA (sa = 1):
 * dE = 2 q (h + dEmat), beta = 1/kT, rank = im.  Uniforms from `words` (the
 * half-step's MT19937 words, 1 per float uniform, 2 per double) or, when
 * words is null, from Philox keyed (key, ctr0, ctr1) at the same position. */
template <class real>
struct BGFlipArgs {
    const real *dEmat;
    int ldE;
    const real *h;
    uint32_t *Q;
    int N, Nw, m;
    int slice0, sliceStep, nEven, sa;
    real twoDivM, coef, beta;
    const uint32_t *words;
    uint32_t key0, key1, ctr0, ctr1;
};
template <class real>
void bipartiteFlipPhase(const BGFlipArgs<real> &a, int nSlices, hipStream_t stream);

/* ---------------- dense brute-force search (dense_bf.hip) ---------------- */

constexpr int kBFMaxLowBits = 12;      /* low bits of x enumerated per block (2^12-entry LDS table) */
constexpr int kBFMaxHighPatterns = 16; /* high patterns per block */

/* Split of x into high and low bits for N (<= 63) */
struct BFPlan {
    int N, L, nH, Llo, lgR, R;
};
BFPlan bfPlan(int N);

/* Device-resident search state: the minimum so far and the list length. */
template <class real>
struct BFState {
    real Emin;
    unsigned long long count;
};

/* Scans x in [xb, xe) (E = x^T W x with W: N x N double, EL: the 2^L-entry
 * low-bit energy table) and merges the range into *d_state and d_xList
 * (capacity cap) with the reference's list rule; asynchronous, no host
 * read-back.  The scratch arrays hold maxBlocks entries; the range may span at
 * most maxBlocks units of 2^(L + lgR) candidates. */
template <class real>
void bfSearchRange(const double *d_W, const double *d_EL, const BFPlan &p, unsigned long long xb,
                   unsigned long long xe, BFState<real> *d_state, unsigned long long cap, real *d_blockMin,
                   unsigned *d_blockCnt, unsigned long long *d_offsets, int maxBlocks,
                   unsigned long long *d_xList, hipStream_t stream);

/* The dense search with the caller's device buffers and a static scratch
 * array (bf_noalloc.hip): W real N x N, results Emin (real, sign of the
 * optimize direction restored), the ascending x list (capacity:
 * min(tileSize rounded up to 256, 2^N)) and its length. */
void denseBFSearchNoalloc(int dtype, const void *d_W, int N, int optimize, unsigned long long tileSize, void *d_Emin,
                          unsigned long long *d_xList, unsigned long long *d_n, hipStream_t stream);

/* ---------------- dense naive step (dense_naive.hip) ---------------- */

template <class real>
struct NaiveArgs {
    int N, Np, m, Nw;
    real twoDivM, coef, beta;          /* host scalars of annealOneStepNaive (:272-273) */
    const real *J;                     /* N x Np rows the fields are updated with */
    const real *h;                     /* Np */
    real *F;                           /* m x Np local fields, in/out */
    uint32_t *Qbits;                   /* m x Nw spins, in/out */
    const uint32_t *words;             /* MT19937 words (3 per float attempt, 4 per double), or null: Philox */
    uint32_t k0, k1, s0, s1;           /* Philox key (seed) and sweep counter */
    unsigned long long *counters;      /* [0] accepted flips, [1] batches */
};
template <class real>
void naiveSweep(const NaiveArgs<real> &a, hipStream_t stream);

/* ---------------- dense sweep (dense_sweep.hip) ---------------- */

template <class real>
struct SweepArgs {
    int N, Np, m, Nw;
    int sa;                 /* 0: SQA coloring, 1: SA (no Trotter coupling) */
    real twoDivM, coef, beta;
    real invKT;
    const real *J;          /* N x Np rows used by the sweep (J, or its bf16 hi+lo rounding) */
    const uint8_t *J8;      /* fp8 form: N x ld8 e4m3 rows, row x scaled by rowScale[x] */
    const float *rowScale;
    int ld8;
    const real *h;          /* Np */
    real *F;                /* m x Np, in/out */
    uint32_t *Qbits;        /* m x Nw, in/out */
    const uint32_t *X;      /* m x N sites */
    const real *U;          /* m x N uniforms */
    uint32_t *D;            /* m x N decisions (tag << 1 | flipped) */
    uint32_t tag;
    int *abortFlag;         /* set by a wave whose wait exceeds the spin limit */
    int *abortHost;         /* host-visible copy of abortFlag (pinned host memory), or null */
    unsigned long long *counters; /* [0] accepted flips, [1] batches, [2] blocked waits */
    int spinLimit;
    int multiBufs;          /* float local-field form: row stage buffers for batches that commit several
                             * flips (>= 2), 0: every batch ends at its first flip */
    int multiMinCand;       /* with multiBufs >= 2: accepting lanes of a batch that make the next batch
                             * commit all its flips */
    int helperWave;         /* float local-field form, single-flip batches, one slice per block: a second
                             * wave per slice applies the deferred J rows to the fields */
};

/* multi_flip: "off", "auto" (the default: multi-flip batches in the first
 * kMultiFlipAutoSweeps sweeps after the spins were set, when most attempts
 * are accepted), "on" (every sweep), each with the threshold "on" stands
 * for unless a number is given. */
enum { kMultiFlipOff = 0, kMultiFlipAuto = 1, kMultiFlipOn = 2 };
constexpr int kMultiFlipDefaultMin = 4;
constexpr int kMultiFlipAutoSweeps = 3;

/* One annealOneStep sweep: N attempts per slice with exact serial-coloring
 * semantics.  plain = true computes each attempt's field as a J-row . q-row
 * dot (the reference algorithm); plain = false keeps the local fields in LDS
 * and updates them with the rows of args.J (fp32 / bf16x2 forms) or of
 * args.J8 (fieldFormat == fieldFP8, float solvers). */
template <class real>
void denseSweep(const SweepArgs<real> &args, bool plain, int fieldFormat, int slicesPerBlock, bool allowDirect,
                int numCUs, hipStream_t stream);

/* Bytes of LDS one block of the local-field sweep needs (ld8: the fp8 row
 * stride, used when fieldFormat == fieldFP8). */
size_t denseSweepLdsBytes(int Np, int slicesPerBlock, size_t realSize, bool plain, int fieldFormat = fieldF32,
                          int ld8 = 0, int multiBufs = 0, bool helper = false);
/* Whether the helper-wave sweep (SweepArgs::helperWave) applies and fits in
 * maxLds bytes per block. */
bool denseSweepHelperFits(int Np, int slicesPerBlock, size_t realSize, bool plain, int fieldFormat,
                          size_t maxLds);
/* Row stage buffers for multi-flip batches that fit beside S slices' regions
 * in maxLds bytes (0: none; at most 4, and (buffers - 1) rows of DMA ops within
 * the 63 outstanding memory operations a wave can count). */
int denseSweepMultiBufs(int Np, int slicesPerBlock, size_t realSize, bool plain, int fieldFormat, size_t maxLds);

/* Many instances in one persistent launch (one wave per (instance, slice)
 * item, taken from a queue in item order; items of one instance must be
 * consecutive and every instance's m must not exceed the resident waves).
 * d_args: one SweepArgs per instance (all of one field format), d_items:
 * (instance, slice) pairs, d_queue: one int, reset by the launcher.
 * regionBytes = denseSweepBatchRegionBytes for the largest instance. */
size_t denseSweepBatchRegionBytes(int maxNp, int maxLd8, size_t realSize, bool plain, int fieldFormat);
/* Blocks of wavesPerBlock waves that stay resident together on the device. */
template <class real>
int denseSweepBatchBlocks(size_t regionBytes, int wavesPerBlock, bool plain, int fieldFormat, int numCUs);
template <class real>
void denseSweepBatch(const SweepArgs<real> *d_args, const uint2 *d_items, int nItems, int *d_queue, bool plain,
                     int fieldFormat, size_t regionBytes, int wavesPerBlock, int nBlocks, hipStream_t stream);

/* The sweep's e4m3 decode applied to the 256 byte codes (a self check). */
void fp8DecodeTable(float out[256]);

/* Row stride of the fp8 rows: Np rounded up to 1 KiB (whole 16-byte-per-lane
 * wave loads). */
inline int fp8RowStride(int Np) { return (Np + 1023) / 1024 * 1024; }

/* ---------------- bipartite brute-force search (bipartite_bf.hip) ---------------- */

constexpr int kBGBFHighPerBlock = 16; /* high x0 patterns per block */

/* One scan rectangle: x1 in [x1b, x1b + nX1), high x0 patterns xH in [hb, he)
 * (x0 = xH << L | xL, all 2^L low patterns); sgn = -1 maximizes. */
struct BGBFArgs {
    int N0, N1, L, nH, K;
    double sgn;
    unsigned long long x1b, nX1, hb, he;
};
BGBFArgs bgBFPlan(int N0, int N1);
unsigned long long bgBFBlocks(const BGBFArgs &a);
/* blockMin[bgBFBlocks(a)]: each block's minimum E */
template <class real>
void bgBFScan(const BGBFArgs &a, const real *d_b0, const real *d_b1, const real *d_W, real *d_blockMin,
              hipStream_t stream);
/* blocks whose minimum is Emin write their first K minimisers (ascending x0,
 * pairs (x0, x1) as 2 u64) to pairs[block * 2K ...] and the count to counts[block] */
template <class real>
void bgBFSelect(const BGBFArgs &a, const real *d_b0, const real *d_b1, const real *d_W, const real *d_blockMin,
                real Emin, unsigned long long *d_pairs, int *d_counts, hipStream_t stream);

/* ---------------- energies (energy.hip) ---------------- */

/* FP64 recompute E[y] = -c - h.q_y - q_y^T J q_y (J N x Np real, symmetric
 * with zero diagonal), spins from packed bits; E returned in double. */
template <class real>
void energyFP64(const real *d_J, const real *d_h, double c, const uint32_t *d_Qbits, int N, int Np, int m, int Nw,
                double *d_partial, double *d_E, hipStream_t stream);

/* Bipartite FP64 recompute E[y] = -c - h0.q0_y - h1.q1_y - q1_y^T J q0_y
 * from JT = J^T (N0 x N1p, zero padded), spins from packed bits (N0p/32 and
 * N1p/32 words per row); d_partial holds m * (N0p + N1p) / 256 doubles. */
template <class real>
void energyBipartiteFP64(const real *d_JT, const real *d_h0, const real *d_h1, double c, const uint32_t *d_Q0,
                         const uint32_t *d_Q1, int N0, int N1, int N0p, int N1p, int m, double *d_partial,
                         double *d_E, hipStream_t stream);

/* Energies on the integer MFMA units (energy_planes.hip).  B (nOut x K,
 * element (i, k) at B[i sRow + k sCol]) is held as P int8 digit planes:
 * B = 2^es sum_p 256^p D_p, exact when `exact` (every element a multiple of
 * 2^es within 8 P - 2 bits), else rounded at 2^es = 2^(max exponent + 3 -
 * 8 P).  finite = 0 if B holds a NaN or an infinity (the planes are then not
 * usable). */
struct EnergyPlaneLayout {
    int nOut = 0, K = 0;
    int outTiles = 0, kBlocks = 0; /* 16-row tiles, 64-column blocks */
    int P = 0, es = 0, exact = 0, finite = 1;
};
/* Scans B and picks the layout; d_scan holds 3 ints; synchronises `stream`. */
template <class real>
EnergyPlaneLayout energyPlaneLayout(const real *d_B, long long sRow, long long sCol, int nOut, int K, int *d_scan,
                                    hipStream_t stream);
/* uint4 elements of the planes; doubles of the partial sums for m slices */
size_t energyPlaneFragCount(const EnergyPlaneLayout &L);
size_t energyPlanePartialCount(const EnergyPlaneLayout &L, int m);
template <class real>
void energyPlaneBuild(const real *d_B, long long sRow, long long sCol, const EnergyPlaneLayout &L, void *d_frag,
                      hipStream_t stream);
/* E[y] = -c - sum_i qout_yi (hOut_i + sum_k B_ik qin_yk) - hIn . qin_y (hIn
 * may be null), spins from packed bits (NwIn, NwOut words per slice row, at
 * least 2 kBlocks and ceil(outTiles / 2) of them). */
template <class real>
void energyPlanes(const EnergyPlaneLayout &L, const void *d_frag, const real *d_hOut, const real *d_hIn, double c,
                  const uint32_t *d_Qin, int NwIn, const uint32_t *d_Qout, int NwOut, int m, double *d_partial,
                  double *d_E, hipStream_t stream);

/* Batched bilinear forms for the formulas objects (energy.hip):
 * E[o] = cst + sgn (lin0 . v + sum_i u_i (lin1_i + sum_j A[i][j] v_j)) with u, v
 * int8 rows (x in {0,1} or q in {-1,+1}) taken from U[iu] and V[iv]: iu = iv =
 * o (twoD = 0) or iu = o / nV, iv = o % nV (twoD = 1, all pairs).  A is
 * nu x nv row-major; lin0 / lin1 may be null.  FP64 accumulation, result
 * rounded to real once. */
template <class real>
struct QuadFormArgs {
    const real *A, *lin0, *lin1;
    double cst, sgn; /* E = cst + sgn * (bilinear + linear terms) */
    const int8_t *U, *V;
    int nu, nv, nV, twoD;
    real *E;
};
template <class real>
void quadForm(const QuadFormArgs<real> &a, int nOut, hipStream_t stream);

/* sum_y q_y . q_(y+1)%m (the Trotter coupling of getSystemE), in double. */
void spinDotSum(const uint32_t *d_Qbits, int N, int m, int Nw, double *d_out, hipStream_t stream);

/* getSystemE of the dense annealer (CPUDenseGraphAnnealer.cpp:230-247) on
 * the device, from the FP64 energies d_E (m) and the slice overlap sum
 * d_spinDot, in real arithmetic in the host formula's order:
 * E = (sum_y sgnE (real)E_y) / m, then E -= (real)spinDot * coef when sqa,
 * then E *= sgnOut; written to *d_out. */
template <class real>
void systemEnergy(const double *d_E, int m, real sgnE, const double *d_spinDot, int sqa, real coef, real sgnOut,
                  real *d_out, hipStream_t stream);

/* ---------------- self checks (diag.hip) ---------------- */

/* sum over float bit patterns b in [lo, hi) of splitmix64((b << 32) | bits(sq_expf(b))),
 * NaN results canonicalised to 0x7fc00000; computed on the current device. */
unsigned long long expfChecksum(unsigned long long lo, unsigned long long hi);

}  // namespace sqaod_rocm
