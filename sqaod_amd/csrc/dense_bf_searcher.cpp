/*
 * DenseGraphBFSearcher<real> of the rocm device: the solver API of
 * sqaodc/cpu/CPUDenseGraphBFSearcher.cpp and sqaodc/cuda/
 * CUDADenseGraphBFSearcher.cpp (state machine Solver.cpp:32-112, tile-size
 * preference Solver.cpp:191-218), with the scan in dense_bf.hip.
 *
 * Result semantics: Emin is the minimum of x^T W x over x in [0, 2^N) (W
 * negated for maximize), and the x list holds the first tile-size x reaching
 * it in ascending order -- the one-worker CPU result, which the CUDA device's
 * per-tile merge also produces.  The list rule is applied on the device per
 * span, so search() runs with no host round trip until makeSolution().
 */
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "solver.h"

namespace sqaod_rocm {

namespace {

/* spans scanned per searchRange call: at least 2^26 candidates (1024 blocks
 * of 65,536), so one call fills the GPU */
constexpr unsigned long long kMinSpan = 1ull << 26;

}  // namespace

template <class real>
void DenseGraphBFSearcher<real>::assignDevice(Device &dev) {
    SQ_THROW_IF(dev_ != nullptr && dev_ != &dev, "Device assigned more than once.");
    dev_ = &dev;
}

template <class real>
void DenseGraphBFSearcher<real>::ensureDevice() {
    if (!dev_)
        dev_ = &defaultDevice();
    dev_->makeCurrent();
}

template <class real>
void DenseGraphBFSearcher<real>::setPreference(const std::string &name, const std::string &value) {
    if (name == "tile_size") {
        /* Solver.cpp:201-209: positive, rounded up to a multiple of 256 */
        long long t = atoll(value.c_str());
        SQ_THROW_IF(t <= 0, "tileSize must be a positive integer.");
        SQ_THROW_IF(t > (1ll << 30), "tileSize %lld is too large.", t);
        tileSize_ = (int)(((t + 255) / 256) * 256);
    } else if (name == "algorithm") {
        /* the searcher has one algorithm; any request selects it
         * (test_dense_graph_bf_searcher.py:132-145) */
    } else if (name == "precision" || name == "device" || name == "experiment") {
        /* read-only / ignored */
    } else {
        SQ_THROW("unknown preference '%s'", name.c_str());
    }
}

template <class real>
PrefList DenseGraphBFSearcher<real>::getPreferences() const {
    /* Solver.cpp:191-198 plus the device entry (CUDADenseGraphBFSearcher.cpp:68-72) */
    PrefList p;
    p.emplace_back("algorithm", algorithmToString(algoBruteForceSearch));
    p.emplace_back("tile_size", std::to_string(tileSize_));
    p.emplace_back("precision", sizeof(real) == 4 ? "float" : "double");
    p.emplace_back("device", "rocm");
    return p;
}

template <class real>
void DenseGraphBFSearcher<real>::setQUBO(const real *W, int N, OptimizeMethod om) {
    /* CUDADenseGraphBFSearcher.cpp:54-67 */
    SQ_THROW_IF(N <= 0, "N must be a positive integer, N=%d.", N);
    SQ_THROW_IF(63 < N, "N must be smaller than 64, N=%d.", N);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < i; ++j)
            SQ_THROW_IF(W[(size_t)i * N + j] != W[(size_t)j * N + i], "W is not symmetric.");
    clearState(solProblemSet);
    N_ = N;
    om_ = om;
    W_.resize((size_t)N * N);
    const double sgn = (om == optMaximize) ? -1. : 1.;
    for (size_t i = 0; i < W_.size(); ++i)
        W_[i] = sgn * (double)W[i]; /* :37-38, exact */
    setState(solProblemSet);
}

template <class real>
void DenseGraphBFSearcher<real>::prepare() {
    /* CPUDenseGraphBFSearcher.cpp:64-89 / CUDADenseGraphBFSearcher.cpp:89-115 */
    throwErrorIfProblemNotSet();
    ensureDevice();
    x_ = 0;
    xMax_ = 1ull << N_;
    if (xMax_ < (unsigned long long)tileSize_) {
        tileSize_ = (int)xMax_; /* the reference keeps the reduced size */
        logf("Tile size is adjusted to %d for N=%d", tileSize_, N_);
    }
    plan_ = bfPlan(N_);
    span_ = std::max((unsigned long long)tileSize_, kMinSpan);
    const unsigned long long unit = 1ull << (plan_.L + plan_.lgR);
    /* a span may straddle unit boundaries at both ends */
    maxBlocks_ = (int)(std::min(span_, xMax_) / unit + 2);

    /* EL[xL] = sum_{b, b'} x_b x_b' W[N-1-b][N-1-b'], double, fixed order */
    const int L = plan_.L, N = N_;
    std::vector<double> EL((size_t)1 << L);
    for (int xL = 0; xL < (1 << L); ++xL) {
        double e = 0.;
        for (int b = 0; b < L; ++b) {
            if (!((xL >> b) & 1))
                continue;
            const double *Wr = W_.data() + (size_t)(N - 1 - b) * N;
            double s = 0.;
            for (int b2 = 0; b2 < L; ++b2)
                if ((xL >> b2) & 1)
                    s += Wr[N - 1 - b2];
            e += s;
        }
        EL[xL] = e;
    }
    d_W_.allocate(W_.size());
    d_EL_.allocate(EL.size());
    d_state_.allocate(1);
    d_blockMin_.allocate((size_t)maxBlocks_);
    d_blockCnt_.allocate((size_t)maxBlocks_);
    d_offsets_.allocate((size_t)maxBlocks_);
    d_xList_.allocate((size_t)tileSize_);
    BFState<real> st;
    st.Emin = (real)FLT_MAX; /* CPUDenseGraphBFSearcher.cpp:65 */
    st.count = 0;
    HIP_CHECK(hipMemcpyAsync(d_W_.ptr, W_.data(), W_.size() * sizeof(double), hipMemcpyHostToDevice, dev_->stream()));
    HIP_CHECK(hipMemcpyAsync(d_EL_.ptr, EL.data(), EL.size() * sizeof(double), hipMemcpyHostToDevice,
                             dev_->stream()));
    HIP_CHECK(hipMemcpyAsync(d_state_.ptr, &st, sizeof(st), hipMemcpyHostToDevice, dev_->stream()));
    dev_->synchronize(); /* the host vectors go out of scope */
    Emin_ = st.Emin;
    xList_.clear();
    E_.clear();
    setState(solPrepared);
}

template <class real>
bool DenseGraphBFSearcher<real>::searchRange(unsigned long long *curXEnd) {
    /* CUDADenseGraphBFSearcher.cpp:155-185, without the per-tile host sync */
    throwErrorIfNotPrepared();
    clearState(solSolutionAvailable);
    dev_->makeCurrent();
    const unsigned long long xb = x_;
    const unsigned long long xe = std::min(x_ + span_, xMax_);
    if (xb < xe)
        bfSearchRange<real>(d_W_.ptr, d_EL_.ptr, plan_, xb, xe, d_state_.ptr, (unsigned long long)tileSize_,
                            d_blockMin_.ptr, d_blockCnt_.ptr, d_offsets_.ptr, maxBlocks_, d_xList_.ptr,
                            dev_->stream());
    x_ = xe;
    if (curXEnd)
        *curXEnd = x_;
    return x_ == xMax_;
}

template <class real>
void DenseGraphBFSearcher<real>::search() {
    /* Solver.cpp:213-218 */
    prepare();
    while (!searchRange(nullptr)) {
    }
    makeSolution();
}

template <class real>
void DenseGraphBFSearcher<real>::makeSolution() {
    /* CUDADenseGraphBFSearcher.cpp:127-150: the device list is already
     * ascending and capped at the tile size */
    throwErrorIfNotPrepared();
    dev_->makeCurrent();
    BFState<real> st;
    HIP_CHECK(hipMemcpyAsync(&st, d_state_.ptr, sizeof(st), hipMemcpyDeviceToHost, dev_->stream()));
    dev_->synchronize();
    const size_t n = (size_t)std::min<unsigned long long>(st.count, (unsigned long long)tileSize_);
    xList_.resize(n);
    if (n) {
        HIP_CHECK(hipMemcpyAsync(xList_.data(), d_xList_.ptr, n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                 dev_->stream()));
        dev_->synchronize();
    }
    Emin_ = st.Emin;
    setState(solSolutionAvailable);
    calculate_E();
}

template <class real>
void DenseGraphBFSearcher<real>::calculate_E() {
    /* CUDADenseGraphBFSearcher.cpp:115-125: E = Emin for every solution (at
     * least one entry), sign restored for maximize */
    throwErrorIfNotPrepared();
    if (!isSolutionAvailable()) {
        dev_->makeCurrent();
        BFState<real> st;
        HIP_CHECK(hipMemcpyAsync(&st, d_state_.ptr, sizeof(st), hipMemcpyDeviceToHost, dev_->stream()));
        dev_->synchronize();
        Emin_ = st.Emin;
    }
    E_.assign(xList_.empty() ? 1 : xList_.size(), om_ == optMaximize ? -Emin_ : Emin_);
    setState(solEAvailable);
}

template <class real>
int DenseGraphBFSearcher<real>::numSolutions() {
    if (!isSolutionAvailable())
        makeSolution();
    return (int)xList_.size();
}

template <class real>
int DenseGraphBFSearcher<real>::numEnergies() {
    if (!isEAvailable())
        calculate_E();
    return (int)E_.size();
}

template <class real>
void DenseGraphBFSearcher<real>::get_x(int8_t *x) {
    if (!isSolutionAvailable())
        makeSolution();
    /* unpackBitSet (Common.cpp:88-93): MSB first */
    for (size_t i = 0; i < xList_.size(); ++i)
        for (int pos = 0; pos < N_; ++pos)
            x[i * N_ + pos] = (int8_t)((xList_[i] >> (N_ - 1 - pos)) & 1ull);
}

template <class real>
void DenseGraphBFSearcher<real>::get_E(real *E) {
    if (!isEAvailable())
        calculate_E();
    memcpy(E, E_.data(), E_.size() * sizeof(real));
}

template class DenseGraphBFSearcher<float>;
template class DenseGraphBFSearcher<double>;

}  // namespace sqaod_rocm
