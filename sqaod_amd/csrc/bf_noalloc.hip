/*
 * The dense brute-force search for callers that cannot allocate device
 * memory (code being captured into a HIP graph, tight serving loops,
 * harnesses that forbid hipMalloc).  It runs the scan of dense_bf.hip -- the
 * same kernels and the same list rule as DenseGraphBFSearcher::search -- but
 * the problem and the results are the caller's device buffers and every
 * intermediate (W in double, the low-bit energy table, the search state and
 * the per-block minima, counts and offsets) lives in `scratch`, a static
 * device array of this library (kScratchBytes per GPU, held until the
 * process ends).  No call allocates or frees memory.
 *
 * Every call on one GPU uses that GPU's one array: calls that overlap on a
 * GPU (two host threads, or two streams whose work overlaps) race.  Each call
 * re-initialises the state it reads (W, the table and the search state are
 * written by the set-up kernel; block entries are written by the scan before
 * they are read), so a call never sees the previous call's contents.
 */
#include <cfloat>
#include <mutex>

#include "common.h"
#include "kernels.h"

namespace sqaod_rocm {

constexpr size_t kScratchBytes = (size_t)256 << 10;

/* the search's intermediates on every GPU the library runs on */
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxN = 63;
constexpr unsigned long long kMaxTile = 1ull << 26;
constexpr unsigned long long kMinSpan = 1ull << 26; /* as the searcher: one call fills the GPU */
constexpr int kMaxBlocks = (int)(kMinSpan >> (kBFMaxLowBits + 4)) + 2;

/* byte offsets inside scratch, each 256-aligned */
constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }
constexpr size_t kOffW = 0;
constexpr size_t kOffEL = align256(kOffW + sizeof(double) * kMaxN * kMaxN);
constexpr size_t kOffState = align256(kOffEL + sizeof(double) * (1u << kBFMaxLowBits));
constexpr size_t kOffBlockMin = align256(kOffState + 64);
constexpr size_t kOffBlockCnt = align256(kOffBlockMin + sizeof(double) * kMaxBlocks);
constexpr size_t kOffOffsets = align256(kOffBlockCnt + sizeof(unsigned) * kMaxBlocks);
constexpr size_t kEnd = align256(kOffOffsets + sizeof(unsigned long long) * kMaxBlocks);
static_assert(kEnd <= kScratchBytes, "scratch layout exceeds the static array");
static_assert(sizeof(BFState<double>) <= 64, "state slot");

std::mutex g_mu;
unsigned char *g_addr[64] = {}; /* scratch's device address per GPU ordinal */

/* The address of this GPU's copy of scratch, looked up once per device with
 * that device current (the array's host-side name is not a device address). */
unsigned char *scratchOnCurrentDevice() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    SQ_THROW_IF(dev < 0 || dev >= 64, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[dev]) {
        void *p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SQ_THROW_IF(p == nullptr, "no device address for the search scratch");
        g_addr[dev] = (unsigned char *)p;
    }
    return g_addr[dev];
}

/* W (real, N x N, the caller's) -> sgn * W in double; the low-bit energy
 * table EL[xL] = sum_{b in xL} sum_{b2 in xL} W[N-1-b][N-1-b2] in the order
 * the searcher's prepare() sums it on the host (b ascending, a partial sum
 * per b); the state reset to (FLT_MAX, 0) (CPUDenseGraphBFSearcher.cpp:65). */
template <class real>
__global__ __launch_bounds__(256) void bfNoallocSetupKernel(const real *W, int N, double sgn, int L, double *dW,
                                                            double *EL, BFState<real> *st) {
    for (int i = threadIdx.x; i < N * N; i += blockDim.x)
        dW[i] = sgn * (double)W[i];
    __syncthreads();
    for (int xL = threadIdx.x; xL < (1 << L); xL += blockDim.x) {
        double e = 0.;
        for (int b = 0; b < L; ++b) {
            if (!((xL >> b) & 1))
                continue;
            const double *Wr = dW + (size_t)(N - 1 - b) * N;
            double s = 0.;
            for (int b2 = 0; b2 < L; ++b2)
                if ((xL >> b2) & 1)
                    s += Wr[N - 1 - b2];
            e += s;
        }
        EL[xL] = e;
    }
    if (threadIdx.x == 0) {
        st->Emin = (real)FLT_MAX;
        st->count = 0;
    }
}

/* the results: Emin with the optimize direction's sign, the list length */
template <class real>
__global__ void bfNoallocResultKernel(const BFState<real> *st, double sgn, unsigned long long cap, real *Emin,
                                      unsigned long long *n) {
    Emin[0] = (real)(sgn * (double)st->Emin);
    n[0] = st->count < cap ? st->count : cap;
}

template <class real>
void searchNoalloc(const real *d_W, int N, int optimize, unsigned long long tileSize, real *d_Emin,
                   unsigned long long *d_xList, unsigned long long *d_n, hipStream_t stream) {
    SQ_THROW_IF(N <= 0 || N > kMaxN, "N must be in 1..%d, N=%d", kMaxN, N);
    SQ_THROW_IF(tileSize == 0 || tileSize > kMaxTile, "tile_size must be in 1..%llu, got %llu", kMaxTile, tileSize);
    SQ_THROW_IF(!d_W || !d_Emin || !d_xList || !d_n, "null device buffer");
    /* the searcher's tile size rules: a multiple of 256 (Solver.cpp:201-209),
     * at most 2^N (CPUDenseGraphBFSearcher.cpp:72-76) */
    const unsigned long long xMax = 1ull << N;
    unsigned long long cap = (tileSize + 255) / 256 * 256;
    if (cap > xMax)
        cap = xMax;
    unsigned char *base = scratchOnCurrentDevice();
    double *dW = (double *)(base + kOffW);
    double *EL = (double *)(base + kOffEL);
    BFState<real> *st = (BFState<real> *)(base + kOffState);
    real *blockMin = (real *)(base + kOffBlockMin);
    unsigned *blockCnt = (unsigned *)(base + kOffBlockCnt);
    unsigned long long *offsets = (unsigned long long *)(base + kOffOffsets);

    const BFPlan plan = bfPlan(N);
    const double sgn = optimize ? -1. : 1.;
    bfNoallocSetupKernel<real><<<1, 256, 0, stream>>>(d_W, N, sgn, plan.L, dW, EL, st);
    HIP_CHECK(hipGetLastError());
    const unsigned long long span = cap > kMinSpan ? cap : kMinSpan;
    for (unsigned long long xb = 0; xb < xMax; xb += span) {
        const unsigned long long xe = (xMax - xb) < span ? xMax : xb + span;
        bfSearchRange<real>(dW, EL, plan, xb, xe, st, cap, blockMin, blockCnt, offsets, kMaxBlocks, d_xList,
                            stream);
    }
    bfNoallocResultKernel<real><<<1, 1, 0, stream>>>(st, sgn, cap, d_Emin, d_n);
    HIP_CHECK(hipGetLastError());
}

}  // namespace

void denseBFSearchNoalloc(int dtype, const void *d_W, int N, int optimize, unsigned long long tileSize, void *d_Emin,
                          unsigned long long *d_xList, unsigned long long *d_n, hipStream_t stream) {
    if (dtype == 0)
        searchNoalloc<float>((const float *)d_W, N, optimize, tileSize, (float *)d_Emin, d_xList, d_n, stream);
    else if (dtype == 1)
        searchNoalloc<double>((const double *)d_W, N, optimize, tileSize, (double *)d_Emin, d_xList, d_n, stream);
    else
        SQ_THROW("dtype must be 0 (float32) or 1 (float64)");
}

}  // namespace sqaod_rocm
