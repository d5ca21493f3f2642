/*
 * BipartiteGraphBFSearcher<real> of the rocm device: the solver API of
 * sqaodc/cpu/CPUBipartiteGraphBFSearcher.cpp and sqaodc/cuda/
 * CUDABipartiteGraphBFSearcher.cpp (state machine Solver.cpp:32-112,
 * tile-size preferences Solver.cpp:221-251), with the scan in
 * bipartite_bf.hip.
 *
 * Result semantics: Emin is the minimum of b0.x0 + b1.x1 + x1^T W x0 over all
 * pairs (b0, b1, W negated for maximize, :39-47), and the list holds the
 * first N0 + N1 pairs reaching it in one worker's visiting order (x0 tile,
 * x1, x0 within the tile; CPUBipartiteGraphBatchSearch.cpp:47-60) -- the
 * one-worker CPU result.  Each searchRange scans a rectangle of x1 values
 * times high-x0 patterns; the device returns each block's first minimisers
 * and the host merges them by visiting key, so the result does not depend on
 * how the pairs were cut into rectangles.
 */
#include <algorithm>
#include <cfloat>
#include <cstring>

#include "solver.h"

namespace sqaod_rocm {

namespace {
constexpr unsigned long long kMaxBlocks = 1ull << 15; /* blocks per rectangle (each 16 x 2^L pairs) */
}  // namespace

template <class real>
void BipartiteGraphBFSearcher<real>::assignDevice(Device &dev) {
    SQ_THROW_IF(dev_ != nullptr && dev_ != &dev, "Device assigned more than once.");
    dev_ = &dev;
}

template <class real>
void BipartiteGraphBFSearcher<real>::ensureDevice() {
    if (!dev_)
        dev_ = &defaultDevice();
    dev_->makeCurrent();
}

template <class real>
void BipartiteGraphBFSearcher<real>::setPreference(const std::string &name, const std::string &value) {
    if (name == "tile_size_0" || name == "tile_size_1") {
        /* Solver.cpp:234-249: positive, rounded up to a multiple of 256 */
        long long t = atoll(value.c_str());
        SQ_THROW_IF(t <= 0, "%s must be a positive integer.", name == "tile_size_0" ? "tileSize0" : "tileSize1");
        SQ_THROW_IF(t > (1ll << 30), "tile size %lld is too large.", t);
        const int r = (int)(((t + 255) / 256) * 256);
        (name == "tile_size_0" ? tileSize0_ : tileSize1_) = r;
    } else if (name == "algorithm" || name == "precision" || name == "device" || name == "experiment") {
        /* the searcher has one algorithm; read-only / ignored names */
    } else {
        SQ_THROW("unknown preference '%s'", name.c_str());
    }
}

template <class real>
PrefList BipartiteGraphBFSearcher<real>::getPreferences() const {
    PrefList p;
    p.emplace_back("algorithm", "brute_force_search");
    p.emplace_back("tile_size_0", std::to_string(tileSize0_));
    p.emplace_back("tile_size_1", std::to_string(tileSize1_));
    p.emplace_back("precision", sizeof(real) == 4 ? "float" : "double");
    p.emplace_back("device", "rocm");
    return p;
}

template <class real>
void BipartiteGraphBFSearcher<real>::setQUBO(const real *b0, const real *b1, const real *W, int N0, int N1,
                                             OptimizeMethod om) {
    /* CPUBipartiteGraphBFSearcher.cpp:30-51 */
    SQ_THROW_IF(N0 <= 0 || N1 <= 0, "b0, b1 must be non-empty.");
    SQ_THROW_IF(63 < N0, "N0 must be smaller than 64, N0=%d.", N0);
    SQ_THROW_IF(63 < N1, "N1 must be smaller than 64, N1=%d.", N1);
    clearState(solProblemSet);
    ensureDevice();
    N0_ = N0;
    N1_ = N1;
    om_ = om;
    d_b0_.allocate((size_t)N0);
    d_b1_.allocate((size_t)N1);
    d_W_.allocate((size_t)N0 * N1);
    HIP_CHECK(hipMemcpy(d_b0_.ptr, b0, (size_t)N0 * sizeof(real), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(d_b1_.ptr, b1, (size_t)N1 * sizeof(real), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(d_W_.ptr, W, (size_t)N0 * N1 * sizeof(real), hipMemcpyHostToDevice));
    setState(solProblemSet);
}

template <class real>
void BipartiteGraphBFSearcher<real>::getProblemSize(int *N0, int *N1) const {
    throwErrorIfProblemNotSet();
    *N0 = N0_;
    *N1 = N1_;
}

template <class real>
unsigned long long BipartiteGraphBFSearcher<real>::visitKey(unsigned long long x0, unsigned long long x1) const {
    return (x0 / T0_) * (T0_ << N1_) + x1 * T0_ + x0 % T0_;
}

template <class real>
void BipartiteGraphBFSearcher<real>::prepare() {
    /* :75-104 */
    throwErrorIfProblemNotSet();
    ensureDevice();
    Emin_ = (real)FLT_MAX;
    list_.clear();
    T0_ = std::min<unsigned long long>((unsigned long long)tileSize0_, 1ull << N0_);
    plan_ = bgBFPlan(N0_, N1_);
    plan_.sgn = (om_ == optMaximize) ? -1. : 1.;
    /* rectangles: up to 4096 x1 values times enough high patterns for kMaxBlocks blocks */
    x1Chunk_ = std::min<unsigned long long>(1ull << N1_, 4096);
    const unsigned long long nHigh = 1ull << plan_.nH;
    hChunk_ = std::min<unsigned long long>(nHigh, std::max<unsigned long long>(1, kMaxBlocks / x1Chunk_) *
                                                      (unsigned long long)kBGBFHighPerBlock);
    hNext_ = 0;
    x1Next_ = 0;
    const unsigned long long blocks = x1Chunk_ * ((hChunk_ + kBGBFHighPerBlock - 1) / kBGBFHighPerBlock);
    d_blockMin_.allocate(blocks);
    d_counts_.allocate(blocks);
    d_pairs_.allocate(blocks * 2 * (size_t)plan_.K);
    setState(solPrepared);
}

template <class real>
bool BipartiteGraphBFSearcher<real>::searchRange(unsigned long long *curX0, unsigned long long *curX1) {
    throwErrorIfNotPrepared();
    clearState(solSolutionAvailable);
    ensureDevice();
    const unsigned long long nHigh = 1ull << plan_.nH, x1max = 1ull << N1_;
    if (hNext_ < nHigh) {
        BGBFArgs a = plan_;
        a.hb = hNext_;
        a.he = std::min(hNext_ + hChunk_, nHigh);
        a.x1b = x1Next_;
        a.nX1 = std::min(x1Chunk_, x1max - x1Next_);
        const unsigned long long blocks = bgBFBlocks(a);
        hipStream_t st = dev_->stream();
        bgBFScan<real>(a, d_b0_.ptr, d_b1_.ptr, d_W_.ptr, d_blockMin_.ptr, st);
        std::vector<real> bm(blocks);
        HIP_CHECK(hipMemcpyAsync(bm.data(), d_blockMin_.ptr, blocks * sizeof(real), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        const real rmin = *std::min_element(bm.begin(), bm.end());
        if (rmin <= Emin_) {
            if (rmin < Emin_) {
                Emin_ = rmin;
                list_.clear();
            }
            bgBFSelect<real>(a, d_b0_.ptr, d_b1_.ptr, d_W_.ptr, d_blockMin_.ptr, rmin, d_pairs_.ptr, d_counts_.ptr,
                             st);
            std::vector<int> cnt(blocks);
            HIP_CHECK(hipMemcpyAsync(cnt.data(), d_counts_.ptr, blocks * sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            const size_t K = (size_t)plan_.K;
            std::vector<unsigned long long> buf(2 * K);
            std::vector<std::pair<unsigned long long, std::pair<unsigned long long, unsigned long long>>> cand;
            for (auto &p : list_)
                cand.push_back({visitKey(p.first, p.second), p});
            for (unsigned long long b = 0; b < blocks; ++b) {
                if (cnt[b] == 0)
                    continue;
                HIP_CHECK(hipMemcpy(buf.data(), d_pairs_.ptr + b * 2 * K, (size_t)cnt[b] * 2 * sizeof(unsigned long long),
                                    hipMemcpyDeviceToHost));
                for (int k = 0; k < cnt[b]; ++k)
                    cand.push_back({visitKey(buf[2 * k], buf[2 * k + 1]), {buf[2 * k], buf[2 * k + 1]}});
            }
            std::sort(cand.begin(), cand.end());
            list_.clear();
            for (size_t k = 0; k < cand.size() && k < K; ++k)
                list_.push_back(cand[k].second);
        }
        /* next rectangle: x1 chunks inside, high-pattern chunks outside */
        x1Next_ += a.nX1;
        if (x1Next_ >= x1max) {
            x1Next_ = 0;
            hNext_ = a.he;
        }
    }
    if (curX0)
        *curX0 = hNext_ << plan_.L;
    if (curX1)
        *curX1 = x1Next_;
    return hNext_ >= nHigh;
}

template <class real>
void BipartiteGraphBFSearcher<real>::search() {
    /* Solver.cpp:253-257 */
    prepare();
    while (!searchRange(nullptr, nullptr))
        ;
    makeSolution();
}

template <class real>
void BipartiteGraphBFSearcher<real>::makeSolution() {
    /* :125-160: the list, E = Emin per pair (sign restored for maximize) */
    throwErrorIfNotPrepared();
    E_.assign(list_.size(), (om_ == optMaximize) ? -Emin_ : Emin_);
    setState(solSolutionAvailable);
    setState(solEAvailable);
}

template <class real>
void BipartiteGraphBFSearcher<real>::calculate_E() {
    /* :113-123: one entry when the list is empty */
    throwErrorIfNotPrepared();
    E_.assign(std::max<size_t>(1, list_.size()), (om_ == optMaximize) ? -Emin_ : Emin_);
    setState(solEAvailable);
}

template <class real>
int BipartiteGraphBFSearcher<real>::numSolutions() {
    if (!isSolutionAvailable())
        makeSolution();
    return (int)list_.size();
}

template <class real>
int BipartiteGraphBFSearcher<real>::numEnergies() {
    if (!isEAvailable())
        calculate_E();
    return std::max<int>(1, (int)E_.size());
}

template <class real>
void BipartiteGraphBFSearcher<real>::get_x(int8_t *x0, int8_t *x1) {
    /* unpackBitSet (Common.cpp:88-93): x[pos] = bit (N-1-pos) */
    if (!isSolutionAvailable())
        makeSolution();
    for (size_t k = 0; k < list_.size(); ++k) {
        for (int pos = 0; pos < N0_; ++pos)
            x0[k * N0_ + pos] = (int8_t)((list_[k].first >> (N0_ - 1 - pos)) & 1ull);
        for (int pos = 0; pos < N1_; ++pos)
            x1[k * N1_ + pos] = (int8_t)((list_[k].second >> (N1_ - 1 - pos)) & 1ull);
    }
}

template <class real>
void BipartiteGraphBFSearcher<real>::get_E(real *E) {
    if (!isEAvailable())
        calculate_E();
    if (E_.empty())
        E[0] = (om_ == optMaximize) ? -Emin_ : Emin_;
    else
        std::copy(E_.begin(), E_.end(), E);
}

template class BipartiteGraphBFSearcher<float>;
template class BipartiteGraphBFSearcher<double>;

}  // namespace sqaod_rocm
