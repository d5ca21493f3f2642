/*
 * BipartiteGraphAnnealer<real> of the rocm device: the solver API of
 * sqaodc/cpu/CPUBipartiteGraphAnnealer.cpp and sqaodc/cuda/
 * CUDABipartiteGraphAnnealer.cu (state machine Solver.cpp:32-112).
 *
 * One annealOneStep is two half-steps (CPUBipartiteGraphAnnealer.cpp:365-373):
 * q1 against the fixed q0, then q0 against the fixed q1.  A half-step is the
 * spin GEMM dEmat = qFixed . J^T (MFMA, dense_field_mfma.hip, or the vector
 * kernel for double solvers) followed by the colour phases of bipartite.hip.
 * Algorithms: coloring and sa_coloring, with the CUDA device's fallbacks
 * (CUDABipartiteGraphAnnealer.cu:102-113); m == 1 selects sa_coloring.
 *
 * Random streams: "philox" (default, counter-based, nothing stored) or
 * "mt19937", the serial sqaod.cpu stream replayed on the device: randomizeSpin
 * draws the q0 row then the q1 row of each slice (:183-192), a half-step
 * draws one uniform per spin in the CPU's visiting order.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <ctime>

#include "solver.h"

namespace sqaod_rocm {

template <class real>
void BipartiteGraphAnnealer<real>::assignDevice(Device &dev) {
    SQ_THROW_IF(dev_ != nullptr && dev_ != &dev, "Device assigned more than once.");
    dev_ = &dev;
}

template <class real>
void BipartiteGraphAnnealer<real>::ensureDevice() {
    if (!dev_)
        dev_ = &defaultDevice();
    dev_->makeCurrent();
}

template <class real>
void BipartiteGraphAnnealer<real>::seed(unsigned long long s) {
    /* :30-35 (worker 0 = s) */
    seed_ = s;
    stepIdx_ = 0;
    mtValid_ = false; /* the device generator restarts from the seed at the next prepare() */
    setState(solRandSeedGiven);
}

template <class real>
Algorithm BipartiteGraphAnnealer<real>::selectAlgorithm(Algorithm algo) {
    /* CUDABipartiteGraphAnnealer.cu:102-113 with Solver.cpp:153-168 */
    if (algo == algoColoring || algo == algoSAColoring) {
        algo_ = algo;
        return algo_;
    }
    Algorithm prev = algo;
    if (algo == algoDefault)
        algo_ = algoColoring;
    else if (algo == algoSADefault)
        algo_ = algoSAColoring;
    else
        algo_ = isSQAAlgorithm(algo) ? algoColoring : algoSAColoring;
    if (algo != algoDefault && algo != algoSADefault)
        logf("%s is not supported, selecting the default algorithm of %s.", algorithmToString(prev),
             algorithmToString(algo_));
    return algo_;
}

template <class real>
void BipartiteGraphAnnealer<real>::setPreference(const std::string &name, const std::string &value) {
    if (name == "n_trotters") {
        /* Solver.cpp:139-146 */
        int m = atoi(value.c_str());
        SQ_THROW_IF(m <= 0, "# trotters must be a positive integer.");
        if (m != m_)
            clearState(solPrepared);
        m_ = m;
    } else if (name == "algorithm") {
        Algorithm a = algorithmFromString(value.c_str());
        if (a == algoUnknown)
            a = (Algorithm)atoi(value.c_str());
        selectAlgorithm(a);
    } else if (name == "rng") {
        if (value == "mt19937")
            rngMT_ = true;
        else if (value == "philox")
            rngMT_ = false;
        else
            SQ_THROW("unknown rng '%s' (philox | mt19937)", value.c_str());
        mtValid_ = false;
        clearState(solPrepared);
    } else if (name == "field_precision") {
        int f;
        if (value == "fp32" || value == "fp64" || value == "native")
            f = fieldF32;
        else if (value == "bf16x2")
            f = fieldBF16x2;
        else
            SQ_THROW("unknown field_precision '%s' (fp32 | bf16x2)", value.c_str());
        SQ_THROW_IF(f == fieldBF16x2 && sizeof(real) != 4, "bf16x2 fields are available for float32 solvers");
        if (f != fieldFormat_)
            fragValid_ = false;
        fieldFormat_ = f;
    } else if (name == "refresh_engine") {
        if (value == "mfma")
            refreshVector_ = false;
        else if (value == "vector")
            refreshVector_ = true;
        else
            SQ_THROW("unknown refresh_engine '%s' (mfma | vector)", value.c_str());
    } else if (name == "precision" || name == "device" || name == "experiment" || name == "tile_size") {
        /* read-only / ignored, as the reference annealers do */
    } else {
        SQ_THROW("unknown preference '%s'", name.c_str());
    }
}

template <class real>
PrefList BipartiteGraphAnnealer<real>::getPreferences() const {
    PrefList p;
    p.emplace_back("algorithm", algorithmToString(algo_));
    p.emplace_back("n_trotters", std::to_string(m_));
    p.emplace_back("precision", sizeof(real) == 4 ? "float" : "double");
    p.emplace_back("device", "rocm");
    p.emplace_back("rng", rngMT_ ? "mt19937" : "philox");
    p.emplace_back("field_precision", fieldFormat_ == fieldBF16x2 ? "bf16x2" : (sizeof(real) == 4 ? "fp32" : "fp64"));
    p.emplace_back("refresh_engine", refreshVector_ ? "vector" : "mfma");
    return p;
}

template <class real>
void BipartiteGraphAnnealer<real>::uploadProblem(const std::vector<double> &h0, const std::vector<double> &h1,
                                                 const std::vector<double> &J, double c) {
    ensureDevice();
    const int N0 = N0_, N1 = N1_;
    N0p_ = roundUp(N0, 256);
    N1p_ = roundUp(N1, 256);
    /* J: N1 x N0p and J^T: N0 x N1p, zero padded, in real */
    std::vector<real> Jp((size_t)N1 * N0p_, real(0)), JTp((size_t)N0 * N1p_, real(0));
    for (int i = 0; i < N1; ++i)
        for (int j = 0; j < N0; ++j) {
            const real v = (real)J[(size_t)i * N0 + j];
            Jp[(size_t)i * N0p_ + j] = v;
            JTp[(size_t)j * N1p_ + i] = v;
        }
    std::vector<real> h0p((size_t)N0p_, real(0)), h1p((size_t)N1p_, real(0));
    h0_.resize(N0);
    h1_.resize(N1);
    for (int j = 0; j < N0; ++j)
        h0p[j] = h0_[j] = (real)h0[j];
    for (int i = 0; i < N1; ++i)
        h1p[i] = h1_[i] = (real)h1[i];
    c_ = (real)c;
    d_J_.allocate(Jp.size());
    d_JT_.allocate(JTp.size());
    d_h0_.allocate(h0p.size());
    d_h1_.allocate(h1p.size());
    HIP_CHECK(hipMemcpyAsync(d_J_.ptr, Jp.data(), Jp.size() * sizeof(real), hipMemcpyHostToDevice, dev_->stream()));
    HIP_CHECK(hipMemcpyAsync(d_JT_.ptr, JTp.data(), JTp.size() * sizeof(real), hipMemcpyHostToDevice, dev_->stream()));
    HIP_CHECK(hipMemcpyAsync(d_h0_.ptr, h0p.data(), h0p.size() * sizeof(real), hipMemcpyHostToDevice, dev_->stream()));
    HIP_CHECK(hipMemcpyAsync(d_h1_.ptr, h1p.data(), h1p.size() * sizeof(real), hipMemcpyHostToDevice, dev_->stream()));
    dev_->synchronize();
    fragValid_ = false;
    planesValid_ = false;
}

template <class real>
void BipartiteGraphAnnealer<real>::hamiltonianOf(const real *b0, const real *b1, const real *W, int N0, int N1,
                                                 OptimizeMethod om, std::vector<double> &h0,
                                                 std::vector<double> &h1, std::vector<double> &J, double &c) {
    /* BGFuncs::calculateHamiltonian (SharedFormulas.cpp:147-167): J = -1/4 W,
     * h0 = -1/4 colsum W - 1/2 b0, h1 = -1/4 rowsum W - 1/2 b1,
     * c = -1/4 sum W - 1/2 (sum b0 + sum b1); all negated for maximize.  Sums
     * are taken in double and rounded once (exact for quarter-integer data). */
    SQ_THROW_IF(N0 <= 0 || N1 <= 0, "N0 and N1 must be positive, N0=%d, N1=%d.", N0, N1);
    const double sgn = (om == optMaximize) ? -1. : 1.;
    h0.assign(N0, 0.);
    h1.assign(N1, 0.);
    J.assign((size_t)N1 * N0, 0.);
    double wsum = 0., b0s = 0., b1s = 0.;
    for (int j = 0; j < N0; ++j) {
        double s = 0.;
        for (int i = 0; i < N1; ++i)
            s += (double)W[(size_t)i * N0 + j];
        h0[j] = sgn * (-0.25 * s - 0.5 * (double)b0[j]);
    }
    for (int i = 0; i < N1; ++i) {
        double s = 0.;
        for (int j = 0; j < N0; ++j) {
            s += (double)W[(size_t)i * N0 + j];
            J[(size_t)i * N0 + j] = sgn * -0.25 * (double)W[(size_t)i * N0 + j];
        }
        wsum += s;
        h1[i] = sgn * (-0.25 * s - 0.5 * (double)b1[i]);
    }
    for (int j = 0; j < N0; ++j)
        b0s += (double)b0[j];
    for (int i = 0; i < N1; ++i)
        b1s += (double)b1[i];
    c = sgn * (-0.25 * wsum - 0.5 * (b0s + b1s));
}

template <class real>
void BipartiteGraphAnnealer<real>::setProblem(const std::vector<double> &h0, const std::vector<double> &h1,
                                              const std::vector<double> &J, double c, int N0, int N1,
                                              OptimizeMethod om) {
    SQ_THROW_IF(N0 <= 0 || N1 <= 0, "N0 and N1 must be positive, N0=%d, N1=%d.", N0, N1);
    clearState(solProblemSet);
    N0_ = N0;
    N1_ = N1;
    m_ = (N0 + N1) / 4; /* :61 */
    om_ = om;
    uploadProblem(h0, h1, J, c);
    setState(solProblemSet);
}

template <class real>
void BipartiteGraphAnnealer<real>::setQUBO(const real *b0, const real *b1, const real *W, int N0, int N1,
                                           OptimizeMethod om) {
    /* :55-75 */
    std::vector<double> h0, h1, J;
    double c;
    hamiltonianOf(b0, b1, W, N0, N1, om, h0, h1, J, c);
    setProblem(h0, h1, J, c, N0, N1, om);
}

template <class real>
void BipartiteGraphAnnealer<real>::setHamiltonian(const real *h0, const real *h1, const real *J, real c, int N0,
                                                  int N1) {
    /* :77-92 */
    SQ_THROW_IF(N0 <= 0 || N1 <= 0, "N0 and N1 must be positive, N0=%d, N1=%d.", N0, N1);
    clearState(solProblemSet);
    N0_ = N0;
    N1_ = N1;
    m_ = (N0 + N1) / 4;
    om_ = optMinimize;
    std::vector<double> vh0(h0, h0 + N0), vh1(h1, h1 + N1), vJ(J, J + (size_t)N0 * N1);
    uploadProblem(vh0, vh1, vJ, (double)c);
    setState(solProblemSet);
}

template <class real>
void BipartiteGraphAnnealer<real>::getHamiltonian(real *h0, real *h1, real *J, real *c) const {
    /* :153-161 */
    throwErrorIfProblemNotSet();
    dev_->makeCurrent();
    memcpy(h0, h0_.data(), (size_t)N0_ * sizeof(real));
    memcpy(h1, h1_.data(), (size_t)N1_ * sizeof(real));
    HIP_CHECK(hipMemcpy2DAsync(J, (size_t)N0_ * sizeof(real), d_J_.ptr, (size_t)N0p_ * sizeof(real),
                               (size_t)N0_ * sizeof(real), N1_, hipMemcpyDeviceToHost, dev_->stream()));
    dev_->synchronize();
    *c = c_;
}

template <class real>
void BipartiteGraphAnnealer<real>::getProblemSize(int *N0, int *N1) const {
    throwErrorIfProblemNotSet();
    *N0 = N0_;
    *N1 = N1_;
}

template <class real>
void BipartiteGraphAnnealer<real>::prepare() {
    /* :225-258 / CUDABipartiteGraphAnnealer.cu:250-292 */
    throwErrorIfProblemNotSet();
    SQ_THROW_IF(m_ <= 0, "# trotters must be a positive integer.");
    ensureDevice();
    if (!isRandSeedGiven())
        seed((unsigned long long)time(NULL));
    setState(solRandSeedGiven);
    if (m_ == 1 && algo_ != algoSAColoring) {
        algo_ = algoSAColoring; /* selectDefaultSAAlgorithm, :234-235 */
        logf("Selecting %s as the default SA algorithm.", algorithmToString(algo_));
    }
    d_Q0_.allocate((size_t)m_ * (N0p_ / 32));
    d_Q1_.allocate((size_t)m_ * (N1p_ / 32));
    HIP_CHECK(hipMemsetAsync(d_Q0_.ptr, 0, d_Q0_.bytes(), dev_->stream()));
    HIP_CHECK(hipMemsetAsync(d_Q1_.ptr, 0, d_Q1_.bytes(), dev_->stream()));
    d_dE_.allocate((size_t)m_ * std::max(N0p_, N1p_));
    d_E_.allocate((size_t)m_);
    ensureEnergyPlanes();
    d_spinDot_.allocate(2);
    if (rngMT_) {
        /* like the CPU's generator, the stream continues across prepare();
         * only seed() restarts it */
        if (!mtValid_) {
            d_mt_.allocate(1);
            MTState st;
            mtSeedHost(&st, seed_);
            HIP_CHECK(hipMemcpyAsync(d_mt_.ptr, &st, sizeof(MTState), hipMemcpyHostToDevice, dev_->stream()));
            mtValid_ = true;
        }
        const size_t wpu = sizeof(real) == 4 ? 1 : 2;
        d_words_.allocate((size_t)m_ * (N0_ + N1_) * wpu);
    }
    stepIdx_ = 0;
    dev_->synchronize();
    setState(solPrepared);
}

template <class real>
void BipartiteGraphAnnealer<real>::randomizeSpin() {
    /* :172-213 (one worker): per slice, the q0 row then the q1 row, randInt(2) */
    throwErrorIfNotPrepared();
    ensureDevice();
    const int N0w = N0p_ / 32, N1w = N1p_ / 32;
    if (rngMT_) {
        const long long stride = (long long)N0_ + N1_;
        mtGenerate(d_mt_.ptr, d_words_.ptr, (size_t)m_ * stride, dev_->stream());
        spinsFromWords(d_words_.ptr, stride, N0_, m_, N0w, d_Q0_.ptr, dev_->stream());
        spinsFromWords(d_words_.ptr + N0_, stride, N1_, m_, N1w, d_Q1_.ptr, dev_->stream());
    } else {
        const unsigned long long salt = 2 * stepIdx_;
        spinsFromPhilox(seed_, 0x40000000ull | salt, N0_, m_, N0w, d_Q0_.ptr, dev_->stream());
        spinsFromPhilox(seed_, 0x40000000ull | (salt + 1), N1_, m_, N1w, d_Q1_.ptr, dev_->stream());
        ++stepIdx_;
    }
    setState(solQSet);
}

template <class real>
void BipartiteGraphAnnealer<real>::set_q(const int8_t *q0, int N0, const int8_t *q1, int N1) {
    /* :109-128: one pair broadcast to every slice */
    throwErrorIfNotPrepared();
    SQ_THROW_IF(N0 != N0_, "Dimension of q0, %d,  should be equal to N0, %d.", N0, N0_);
    SQ_THROW_IF(N1 != N1_, "Dimension of q1, %d,  should be equal to N1, %d.", N1, N1_);
    std::vector<int8_t> a0((size_t)m_ * N0_), a1((size_t)m_ * N1_);
    for (int y = 0; y < m_; ++y) {
        memcpy(&a0[(size_t)y * N0_], q0, (size_t)N0_);
        memcpy(&a1[(size_t)y * N1_], q1, (size_t)N1_);
    }
    set_qset(a0.data(), a1.data(), m_, N0, N1);
}

template <class real>
void BipartiteGraphAnnealer<real>::set_qset(const int8_t *q0, const int8_t *q1, int m, int N0, int N1) {
    /* :130-143: m = number of pairs, prepare(), copy */
    throwErrorIfProblemNotSet();
    SQ_THROW_IF(N0 != N0_ || N1 != N1_, "Dimension of q0/q1 (%d, %d) should be (%d, %d).", N0, N1, N0_, N1_);
    SQ_THROW_IF(m <= 0, "empty q set");
    m_ = m;
    prepare();
    d_q8_.allocate((size_t)m_ * std::max(N0_, N1_));
    HIP_CHECK(hipMemcpyAsync(d_q8_.ptr, q0, (size_t)m_ * N0_, hipMemcpyHostToDevice, dev_->stream()));
    packSpins(d_q8_.ptr, N0_, m_, N0p_ / 32, d_Q0_.ptr, dev_->stream());
    dev_->synchronize();
    HIP_CHECK(hipMemcpyAsync(d_q8_.ptr, q1, (size_t)m_ * N1_, hipMemcpyHostToDevice, dev_->stream()));
    packSpins(d_q8_.ptr, N1_, m_, N1p_ / 32, d_Q1_.ptr, dev_->stream());
    dev_->synchronize();
    setState(solQSet);
}

template <class real>
void BipartiteGraphAnnealer<real>::computeFields(int side) {
    /* side 1: dE1 = Q0 . J^T (m x N1p): rows of B are J's rows (N1 of length N0p);
     * side 0: dE0 = Q1 . J   (m x N0p): rows of B are J^T's rows (N0 of length N1p) */
    const int nOut = side ? N1_ : N0_, nOutP = side ? N1p_ : N0p_, K = side ? N0p_ : N1p_;
    const int nIn = side ? N0_ : N1_;
    const uint32_t *Qin = side ? d_Q0_.ptr : d_Q1_.ptr;
    const int NwIn = K / 32;
    const int ldE = std::max(N0p_, N1p_);
    if (sizeof(real) == 4 && !refreshVector_) {
        if (!fragValid_) {
            const float *J = (const float *)d_J_.ptr, *JT = (const float *)d_JT_.ptr;
            if (fieldFormat_ == fieldBF16x2) {
                d_fragJ32_.release();
                d_fragJT32_.release();
                d_fragJ16_.allocate(fieldFragElemsBF16x2(N1p_, N0p_));
                d_fragJT16_.allocate(fieldFragElemsBF16x2(N0p_, N1p_));
                buildFragBF16x2(J, N1_, N0p_, N1p_, N0p_, d_fragJ16_.ptr, d_fragJ16_.n, dev_->stream());
                buildFragBF16x2(JT, N0_, N1p_, N0p_, N1p_, d_fragJT16_.ptr, d_fragJT16_.n, dev_->stream());
            } else {
                d_fragJ16_.release();
                d_fragJT16_.release();
                d_fragJ32_.allocate(fieldFragElemsF32(N1p_, N0p_));
                d_fragJT32_.allocate(fieldFragElemsF32(N0p_, N1p_));
                buildFragF32(J, N1_, N0p_, N1p_, N0p_, d_fragJ32_.ptr, d_fragJ32_.n, dev_->stream());
                buildFragF32(JT, N0_, N1p_, N0p_, N1p_, d_fragJT32_.ptr, d_fragJT32_.n, dev_->stream());
            }
            fragValid_ = true;
        }
        if (fieldFormat_ == fieldBF16x2) {
            const DevBuf<uint16_t> &f = side ? d_fragJ16_ : d_fragJT16_;
            fieldRefreshBF16x2(f.ptr, f.n, Qin, nOutP, K, m_, NwIn, ldE, (float *)d_dE_.ptr, dev_->stream());
        } else {
            const DevBuf<uint16_t> &f = side ? d_fragJ32_ : d_fragJT32_;
            fieldRefreshF32MFMA(f.ptr, f.n, Qin, nOutP, K, m_, NwIn, ldE, (float *)d_dE_.ptr, dev_->stream());
        }
    } else {
        /* vector form: dE[y][o] = sum_{k < nIn} q[y][k] B[k][o] with B = J^T (side 1) or J (side 0) */
        const real *B = side ? d_JT_.ptr : d_J_.ptr;
        fieldRefresh<real>(B, nIn, nOutP, nOutP, Qin, m_, NwIn, ldE, d_dE_.ptr, dev_->stream());
    }
    (void)nOut;
}

template <class real>
void BipartiteGraphAnnealer<real>::halfStep(int side, real G, real beta, const uint32_t *words) {
    /* annealHalfStepColoring (:346-362) / annealHalfStepSAColoring (:498-530) */
    computeFields(side);
    BGFlipArgs<real> a{};
    a.dEmat = d_dE_.ptr;
    a.ldE = std::max(N0p_, N1p_);
    a.h = side ? d_h1_.ptr : d_h0_.ptr;
    a.Q = side ? d_Q1_.ptr : d_Q0_.ptr;
    a.N = side ? N1_ : N0_;
    a.Nw = (side ? N1p_ : N0p_) / 32;
    a.m = m_;
    a.nEven = (m_ + 1) / 2;
    a.words = words;
    a.key0 = (uint32_t)seed_;
    a.key1 = (uint32_t)(seed_ >> 32);
    a.ctr0 = (uint32_t)stepIdx_;
    a.ctr1 = (uint32_t)(stepIdx_ >> 32) ^ (side ? 0x80000000u : 0x40000000u);
    if (algo_ == algoSAColoring) {
        a.sa = 1;
        a.beta = real(1.) / G; /* annealOneStepSAColoring: invKT = 1 / kT, kT passed as G */
        a.slice0 = 0;
        a.sliceStep = 1;
        bipartiteFlipPhase<real>(a, m_, dev_->stream());
        return;
    }
    a.sa = 0;
    a.twoDivM = real(2.) / real(m_);
    a.coef = std::log(std::tanh(G * beta / real(m_))) / beta;
    a.beta = beta;
    a.sliceStep = 2;
    /* even slices (all of them when m is even; without m-1 when m is odd) */
    const int nEvenFirst = (m_ % 2 == 0) ? m_ / 2 : (m_ - 1) / 2;
    a.slice0 = 0;
    bipartiteFlipPhase<real>(a, nEvenFirst, dev_->stream());
    if (m_ % 2 == 1) {
        /* slice m-1: even, adjacent to slice 0, visited after it */
        a.slice0 = m_ - 1;
        bipartiteFlipPhase<real>(a, 1, dev_->stream());
    }
    a.slice0 = 1;
    bipartiteFlipPhase<real>(a, m_ / 2, dev_->stream());
}

template <class real>
void BipartiteGraphAnnealer<real>::annealOneStep(real G, real beta) {
    /* annealOneStepColoring, :365-373 / annealOneStepSAColoring, :533-542 */
    throwErrorIfQNotSet();
    clearState(solSolutionAvailable);
    ensureDevice();
    const uint32_t *w1 = nullptr, *w0 = nullptr;
    if (rngMT_) {
        const size_t wpu = sizeof(real) == 4 ? 1 : 2;
        const size_t n1 = (size_t)m_ * N1_ * wpu, n0 = (size_t)m_ * N0_ * wpu;
        mtGenerate(d_mt_.ptr, d_words_.ptr, n1 + n0, dev_->stream());
        w1 = d_words_.ptr;
        w0 = d_words_.ptr + n1;
    }
    halfStep(1, G, beta, w1);
    halfStep(0, G, beta, w0);
    ++stepIdx_;
}

template <class real>
void BipartiteGraphAnnealer<real>::ensureEnergyPlanes() {
    /* J (N1 x N0, row stride N0p) as digit planes, rows for q1 and columns
     * for q0; synchronises the stream when the planes are built */
    if (!planesValid_) {
        d_planeScan_.allocate(3);
        planes_ = energyPlaneLayout<real>(d_J_.ptr, N0p_, 1, N1_, N0_, d_planeScan_.ptr, dev_->stream());
        if (planes_.finite) {
            d_planes_.allocate(energyPlaneFragCount(planes_));
            energyPlaneBuild<real>(d_J_.ptr, N0p_, 1, planes_, d_planes_.ptr, dev_->stream());
        } else {
            d_planes_.release();
        }
        planesValid_ = true;
    }
    size_t parts = (size_t)m_ * ((N0p_ + N1p_) / 256); /* the FP64 vector kernel's */
    if (planes_.finite)
        parts = std::max(parts, energyPlanePartialCount(planes_, m_));
    d_partial_.allocate(parts);
}

template <class real>
void BipartiteGraphAnnealer<real>::computeEnergies() {
    ensureEnergyPlanes();
    if (planes_.finite)
        energyPlanes<real>(planes_, d_planes_.ptr, d_h1_.ptr, d_h0_.ptr, (double)c_, d_Q0_.ptr, N0p_ / 32, d_Q1_.ptr,
                           N1p_ / 32, m_, d_partial_.ptr, d_E_.ptr, dev_->stream());
    else
        energyBipartiteFP64<real>(d_JT_.ptr, d_h0_.ptr, d_h1_.ptr, (double)c_, d_Q0_.ptr, d_Q1_.ptr, N0_, N1_, N0p_,
                                  N1p_, m_, d_partial_.ptr, d_E_.ptr, dev_->stream());
}

template <class real>
void BipartiteGraphAnnealer<real>::calculate_E() {
    /* :215-222 with BGFuncs::calculate_E (SharedFormulas.cpp:184-202),
     * evaluated exactly (integer planes) or to FP64 accuracy */
    throwErrorIfQNotSet();
    ensureDevice();
    computeEnergies();
    std::vector<double> e((size_t)m_);
    HIP_CHECK(hipMemcpyAsync(e.data(), d_E_.ptr, (size_t)m_ * sizeof(double), hipMemcpyDeviceToHost, dev_->stream()));
    dev_->synchronize();
    const real sgn = (om_ == optMaximize) ? real(-1.) : real(1.);
    E_.resize((size_t)m_);
    Efp64_.resize((size_t)m_);
    for (int y = 0; y < m_; ++y) {
        E_[y] = sgn * (real)e[y];
        Efp64_[y] = (double)sgn * e[y];
    }
    setState(solEAvailable);
}

template <class real>
void BipartiteGraphAnnealer<real>::makeSolution() {
    /* :260-266: syncBits then calculate_E */
    throwErrorIfNotPrepared();
    throwErrorIfQNotSet();
    ensureDevice();
    d_q8_.allocate((size_t)m_ * std::max(N0_, N1_));
    q0_.resize((size_t)m_ * N0_);
    q1_.resize((size_t)m_ * N1_);
    unpackSpins(d_Q0_.ptr, N0_, m_, N0p_ / 32, d_q8_.ptr, dev_->stream());
    HIP_CHECK(hipMemcpyAsync(q0_.data(), d_q8_.ptr, q0_.size(), hipMemcpyDeviceToHost, dev_->stream()));
    dev_->synchronize();
    unpackSpins(d_Q1_.ptr, N1_, m_, N1p_ / 32, d_q8_.ptr, dev_->stream());
    HIP_CHECK(hipMemcpyAsync(q1_.data(), d_q8_.ptr, q1_.size(), hipMemcpyDeviceToHost, dev_->stream()));
    dev_->synchronize();
    setState(solSolutionAvailable);
    calculate_E();
}

template <class real>
void BipartiteGraphAnnealer<real>::get_q(int8_t *q0, int8_t *q1) {
    if (!isSolutionAvailable())
        makeSolution();
    memcpy(q0, q0_.data(), q0_.size());
    memcpy(q1, q1_.data(), q1_.size());
}

template <class real>
void BipartiteGraphAnnealer<real>::get_x(int8_t *x0, int8_t *x1) {
    /* x = (q + 1) / 2 (:546-557, Common.cpp:133-145) */
    if (!isSolutionAvailable())
        makeSolution();
    for (size_t i = 0; i < q0_.size(); ++i)
        x0[i] = (int8_t)((q0_[i] + 1) / 2);
    for (size_t i = 0; i < q1_.size(); ++i)
        x1[i] = (int8_t)((q1_[i] + 1) / 2);
}

template <class real>
void BipartiteGraphAnnealer<real>::get_E(real *E) {
    if (!isEAvailable())
        calculate_E();
    memcpy(E, E_.data(), (size_t)m_ * sizeof(real));
}

template <class real>
real BipartiteGraphAnnealer<real>::getSystemE(real G, real beta) {
    /* :268-286 */
    calculate_E();
    real E = real(0.);
    for (int y = 0; y < m_; ++y)
        E += E_[y];
    E = E / real(m_);
    if (isSQAAlgorithm(algo_)) {
        double dots[2] = {0., 0.};
        spinDotSum(d_Q0_.ptr, N0_, m_, N0p_ / 32, d_spinDot_.ptr, dev_->stream());
        spinDotSum(d_Q1_.ptr, N1_, m_, N1p_ / 32, d_spinDot_.ptr + 1, dev_->stream());
        HIP_CHECK(hipMemcpyAsync(dots, d_spinDot_.ptr, 2 * sizeof(double), hipMemcpyDeviceToHost, dev_->stream()));
        dev_->synchronize();
        real spinDotSum_ = (real)(dots[0] + dots[1]);
        real coef = real(0.5) / beta * std::log(std::tanh(G * beta / real(m_)));
        E -= spinDotSum_ * coef;
    }
    if (om_ == optMaximize)
        E *= real(-1.);
    return E;
}

template class BipartiteGraphAnnealer<float>;
template class BipartiteGraphAnnealer<double>;

}  // namespace sqaod_rocm
