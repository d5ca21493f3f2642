/*
 * rocm DenseGraphAnnealer<real>: the sqaodc dense annealer API
 * (sqaodc/common/Solver.h:187-205) with device-resident spins, local fields
 * and random streams.  Method semantics follow
 * sqaodc/cpu/CPUDenseGraphAnnealer.cpp and sqaodc/cuda/CUDADenseGraphAnnealer.cu;
 * the file:line of the behaviour each method mirrors is given at the method.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <mutex>
#include <strings.h>

#include "solver.h"

namespace sqaod_rocm {

/* ---------------- shared helpers ---------------- */

void logf(const char *fmt, ...) {
    static int verbose = -1;
    if (verbose < 0) {
        const char *v = getenv("SQAOD_VERBOSE");
        verbose = (v && atoi(v) != 0) ? 1 : 0;
    }
    if (!verbose)
        return;
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
}

const char *algorithmToString(Algorithm a) {
    /* Preference.cpp:20-40 */
    switch (a) {
    case algoNaive: return "naive";
    case algoColoring: return "coloring";
    case algoBruteForceSearch: return "brute_force_search";
    case algoSADefault: return "sa_default";
    case algoSANaive: return "sa_naive";
    case algoSAColoring: return "sa_coloring";
    case algoDefault: return "default";
    default: return "unknown";
    }
}

Algorithm algorithmFromString(const char *s) {
    /* Preference.cpp:42-58 */
    if (!strcasecmp("naive", s)) return algoNaive;
    if (!strcasecmp("coloring", s)) return algoColoring;
    if (!strcasecmp("brute_force_search", s)) return algoBruteForceSearch;
    if (!strcasecmp("sa_default", s)) return algoSADefault;
    if (!strcasecmp("sa_naive", s)) return algoSANaive;
    if (!strcasecmp("sa_coloring", s)) return algoSAColoring;
    if (!strcasecmp("default", s)) return algoDefault;
    return algoUnknown;
}

bool isSQAAlgorithm(Algorithm a) { return a == algoNaive || a == algoColoring; }

void SolverState::clearState(int s) {
    /* Solver.cpp:39-64: clearing a state clears every state that depends on it */
    int toClear = 0;
    switch (s) {
    case solRandSeedGiven:
        toClear = solRandSeedGiven;
        break;
    case solProblemSet:
        toClear |= solProblemSet;
        /* fall through */
    case solPrepared:
        toClear |= solPrepared;
        /* fall through */
    case solQSet:
        toClear |= solQSet;
        /* fall through */
    case solEAvailable:
    case solSolutionAvailable:
        toClear |= solEAvailable | solSolutionAvailable;
        break;
    default:
        break;
    }
    state_ &= ~toClear;
}

void SolverState::throwErrorIfProblemNotSet() const { SQ_THROW_IF(!isProblemSet(), "Problem is not set."); }

void SolverState::throwErrorIfNotPrepared() const {
    throwErrorIfProblemNotSet();
    SQ_THROW_IF(!isPrepared(), "not prepared, call prepare() in advance.");
}

void SolverState::throwErrorIfQNotSet() const {
    SQ_THROW_IF(!isQSet(), "Bits(x or q) not initialized.  Plase set or randomize in advance.");
}

Device::Device(int devno) : devno_(devno) {
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    SQ_THROW_IF(devno < 0 || devno >= n, "device %d not available (%d devices)", devno, n);
    HIP_CHECK(hipSetDevice(devno));
    HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, devno));
    numCUs_ = prop.multiProcessorCount;
    maxLds_ = prop.sharedMemPerBlock;
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, devno) == hipSuccess && optin > 0)
        maxLds_ = (size_t)optin;
}

namespace {
std::mutex g_sweepersMu;
std::map<int, int> g_sweepers; /* GPU -> Devices that launch sweeps */
}  // namespace

bool Device::soleSweeper() {
    std::lock_guard<std::mutex> lock(g_sweepersMu);
    if (!sweeper_) {
        sweeper_ = true;
        ++g_sweepers[devno_];
    }
    return g_sweepers[devno_] == 1;
}

Device::~Device() {
    if (sweeper_) {
        std::lock_guard<std::mutex> lock(g_sweepersMu);
        --g_sweepers[devno_];
    }
    (void)hipSetDevice(devno_);
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
}

void Device::makeCurrent() const { HIP_CHECK(hipSetDevice(devno_)); }
void Device::synchronize() const { HIP_CHECK(hipStreamSynchronize(stream_)); }

Device &defaultDevice() {
    static Device *dev = new Device(0); /* lives until process exit */
    return *dev;
}

/* ---------------- DenseGraphAnnealer ---------------- */

template <class real>
DenseGraphAnnealer<real>::DenseGraphAnnealer() {
    selectAlgorithm(algoDefault); /* CPUDenseGraphAnnealer.cpp:14-21 */
}

template <class real>
DenseGraphAnnealer<real>::~DenseGraphAnnealer() {
    if (dev_) {
        (void)hipSetDevice(dev_->devno());
        (void)hipStreamSynchronize(dev_->stream());
    }
    for (auto &pe : pending_) {
        (void)hipEventDestroy(pe.start);
        (void)hipEventDestroy(pe.stop);
    }
    for (auto e : eventPool_)
        (void)hipEventDestroy(e);
}

template <class real>
hipEvent_t DenseGraphAnnealer<real>::takeEvent() {
    if (!eventPool_.empty()) {
        hipEvent_t e = eventPool_.back();
        eventPool_.pop_back();
        return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    return e;
}

template <class real>
void DenseGraphAnnealer<real>::phaseBegin(int phase, PhaseEvents &pe) {
    pe.phase = phase;
    pe.start = takeEvent();
    pe.stop = takeEvent();
    HIP_CHECK(hipEventRecord(pe.start, dev_->stream()));
}

template <class real>
void DenseGraphAnnealer<real>::phaseEnd(PhaseEvents &pe) {
    HIP_CHECK(hipEventRecord(pe.stop, dev_->stream()));
    pending_.push_back(pe);
}

template <class real>
void DenseGraphAnnealer<real>::collectPhases() {
    for (auto &pe : pending_) {
        HIP_CHECK(hipEventSynchronize(pe.stop));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, pe.start, pe.stop));
        phaseMs_[pe.phase] += ms;
        phaseCount_[pe.phase] += 1;
        eventPool_.push_back(pe.start);
        eventPool_.push_back(pe.stop);
    }
    pending_.clear();
}

template <class real>
void DenseGraphAnnealer<real>::getPhaseTimes(double ms[3], unsigned long long counts[3]) {
    if (dev_)
        dev_->makeCurrent();
    collectPhases();
    for (int i = 0; i < 3; ++i) {
        ms[i] = phaseMs_[i];
        counts[i] = phaseCount_[i];
        phaseMs_[i] = 0.;
        phaseCount_[i] = 0;
    }
}

template <class real>
void DenseGraphAnnealer<real>::assignDevice(Device &dev) {
    /* cuda/api.h DeviceAssigner: a device is assigned once */
    SQ_THROW_IF(dev_ != nullptr && dev_ != &dev, "Device assigned more than once.");
    dev_ = &dev;
}

template <class real>
void DenseGraphAnnealer<real>::ensureDevice() {
    if (!dev_)
        dev_ = &defaultDevice();
    dev_->makeCurrent();
}

template <class real>
void DenseGraphAnnealer<real>::seed(unsigned long long s) {
    /* :29-33 (the device streams are keyed by the seed; worker 0 = s) */
    seed_ = s;
    sweepIdx_ = 0;
    if (d_mt_.ptr) {
        ensureDevice();
        seedWorkerGenerators();
    }
    setState(solRandSeedGiven);
}

template <class real>
void DenseGraphAnnealer<real>::seedWorkerGenerators() {
    /* worker w of the CPU annealer is seeded s + 17 w (:29-33) */
    std::vector<MTState> st((size_t)nWorkers_);
    for (int w = 0; w < nWorkers_; ++w)
        mtSeedHost(&st[w], seed_ + 17ull * (unsigned long long)w);
    d_mt_.allocate((size_t)nWorkers_);
    HIP_CHECK(hipMemcpyAsync(d_mt_.ptr, st.data(), st.size() * sizeof(MTState), hipMemcpyHostToDevice,
                             dev_->stream()));
    dev_->synchronize();
    mtWorkersSeeded_ = nWorkers_;
}

template <class real>
Algorithm DenseGraphAnnealer<real>::selectAlgorithm(Algorithm algo) {
    /* CPUDenseGraphAnnealer.cpp:35-50: naive, coloring and sa_naive are
     * offered; other requests map to the default of their kind
     * (Solver.cpp:152-168). */
    switch (algo) {
    case algoNaive:
    case algoColoring:
    case algoSANaive:
        algo_ = algo;
        break;
    default: {
        Algorithm prev = algo;
        if (algo == algoDefault)
            algo_ = algoColoring;
        else if (algo == algoSADefault)
            algo_ = algoSANaive;
        else
            algo_ = isSQAAlgorithm(algo) ? algoColoring : algoSANaive;
        if (algo != algoDefault && algo != algoSADefault)
            logf("%s is not supported, selecting the default algorithm of %s.", algorithmToString(prev),
                 algorithmToString(algo_));
        break;
    }
    }
    return algo_;
}

template <class real>
void DenseGraphAnnealer<real>::setPreference(const std::string &name, const std::string &value) {
    /* Solver.cpp:139-150 for n_trotters/algorithm; the rest are rocm-device knobs */
    if (name == "n_trotters") {
        int m = atoi(value.c_str());
        SQ_THROW_IF(m <= 0, "# trotters must be a positive integer.");
        if (m_ != m)
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
            SQ_THROW("unknown rng '%s' (mt19937 | philox)", value.c_str());
        clearState(solPrepared);
    } else if (name == "sweep") {
        if (value == "plain")
            plainSweep_ = true;
        else if (value == "local")
            plainSweep_ = false;
        else
            SQ_THROW("unknown sweep '%s' (local | plain)", value.c_str());
        fieldsDirty_ = true;
    } else if (name == "field_precision") {
        int f;
        if (value == "fp32" || value == "fp64" || value == "native")
            f = fieldF32;
        else if (value == "bf16x2")
            f = fieldBF16x2;
        else if (value == "fp8")
            f = fieldFP8;
        else
            SQ_THROW("unknown field_precision '%s' (fp32 | bf16x2 | fp8)", value.c_str());
        SQ_THROW_IF(f != fieldF32 && sizeof(real) != 4, "%s fields are available for float32 solvers",
                    value.c_str());
        if (f != fieldFormat_) {
            fieldFormat_ = f;
            fieldsDirty_ = true;
            fragValid_ = false;
            if (d_J_.ptr) {
                ensureDevice();
                prepareFieldForm();
            }
        }
    } else if (name == "refresh_engine") {
        if (value == "mfma")
            refreshVector_ = false;
        else if (value == "vector")
            refreshVector_ = true;
        else
            SQ_THROW("unknown refresh_engine '%s' (mfma | vector)", value.c_str());
        fieldsDirty_ = true;
    } else if (name == "slices_per_block") {
        slicesPerBlock_ = atoi(value.c_str());
    } else if (name == "sweep_launch") {
        if (value == "cooperative")
            coopLaunch_ = true;
        else if (value == "auto")
            coopLaunch_ = false;
        else
            SQ_THROW("unknown sweep_launch '%s' (auto | cooperative)", value.c_str());
    } else if (name == "refresh_interval") {
        refreshInterval_ = atoi(value.c_str());
    } else if (name == "profile") {
        /* 0: off; 1: HIP events around each sweep; 2: around every phase */
        if (value == "on" || value == "true" || value == "True")
            profile_ = 1;
        else
            profile_ = atoi(value.c_str()) < 0 ? 0 : atoi(value.c_str());
    } else if (name == "n_workers") {
        /* with rng=mt19937: the number of CPU workers (OpenMP threads) of the
         * sqaod.cpu run whose coloring trajectory is replayed */
        const int v = atoi(value.c_str());
        SQ_THROW_IF(v < 1 || v > 1024, "n_workers must be in 1..1024");
        if (v != nWorkers_) {
            nWorkers_ = v;
            clearState(solPrepared);
        }
    } else if (name == "multi_flip") {
        /* float local-field sweeps: "off" (every batch ends at its first
         * flip), "auto" (multi-flip batches in the first sweeps after the
         * spins were set), "on" (in every sweep), or a number 2..64: "on"
         * with that threshold of accepting lanes */
        if (value == "on" || value == "true" || value == "True") {
            multiFlipMode_ = kMultiFlipOn;
            multiFlipMin_ = kMultiFlipDefaultMin;
        } else if (value == "auto") {
            multiFlipMode_ = kMultiFlipAuto;
            multiFlipMin_ = kMultiFlipDefaultMin;
        } else if (value == "off" || value == "0" || value == "false" || value == "False") {
            multiFlipMode_ = kMultiFlipOff;
        } else {
            const int v = atoi(value.c_str());
            SQ_THROW_IF(v < 2 || v > 64 || std::to_string(v) != value, "multi_flip must be auto, on, off or 2..64");
            multiFlipMode_ = kMultiFlipOn;
            multiFlipMin_ = v;
        }
    } else if (name == "helper_wave") {
        /* "on" (default): single-flip float local-field sweeps with one slice
         * per block run a second wave per slice that applies the deferred J
         * rows to the fields; "off": the sweeping wave applies them */
        if (value == "on" || value == "true" || value == "True" || value == "1")
            helperWave_ = true;
        else if (value == "off" || value == "false" || value == "False" || value == "0")
            helperWave_ = false;
        else
            SQ_THROW("helper_wave must be on or off");
    } else if (name == "spin_limit") {
        const int v = atoi(value.c_str());
        SQ_THROW_IF(v <= 0, "spin_limit must be positive");
        spinLimit_ = v;
    } else if (name == "precision" || name == "device") {
        /* read-only (Preference.h:26-36) */
    } else if (name == "experiment" || name == "tile_size") {
        /* accepted and ignored, as the reference annealers do */
    } else {
        SQ_THROW("unknown preference '%s'", name.c_str());
    }
}

template <class real>
PrefList DenseGraphAnnealer<real>::getPreferences() const {
    /* Solver.cpp:130-137 plus the device entry (CUDADenseGraphAnnealer.cu:179-183) */
    PrefList p;
    p.emplace_back("algorithm", algorithmToString(algo_));
    p.emplace_back("n_trotters", std::to_string(m_));
    p.emplace_back("precision", sizeof(real) == 4 ? "float" : "double");
    p.emplace_back("device", "rocm");
    p.emplace_back("rng", rngMT_ ? "mt19937" : "philox");
    p.emplace_back("n_workers", std::to_string(nWorkers_));
    p.emplace_back("sweep", plainSweep_ ? "plain" : "local");
    p.emplace_back("field_precision", fieldFormat_ == fieldBF16x2 ? "bf16x2"
                                      : fieldFormat_ == fieldFP8  ? "fp8"
                                                                  : (sizeof(real) == 4 ? "fp32" : "fp64"));
    p.emplace_back("slices_per_block", std::to_string(slicesPerBlock_));
    p.emplace_back("sweep_launch", coopLaunch_ ? "cooperative" : "auto");
    p.emplace_back("refresh_interval", std::to_string(refreshInterval_));
    p.emplace_back("refresh_engine", refreshVector_ ? "vector" : "mfma");
    p.emplace_back("profile", std::to_string(profile_));
    p.emplace_back("multi_flip", multiFlipMode_ == kMultiFlipOff            ? std::string("off")
                                 : multiFlipMode_ == kMultiFlipAuto         ? std::string("auto")
                                 : multiFlipMin_ == kMultiFlipDefaultMin    ? std::string("on")
                                                                            : std::to_string(multiFlipMin_));
    p.emplace_back("helper_wave", helperWave_ ? "on" : "off");
    return p;
}

template <class real>
void DenseGraphAnnealer<real>::allocateProblem(int N) {
    ensureDevice();
    N_ = N;
    Np_ = roundUp(N, 256);
    Nw_ = Np_ / 32;
    m_ = N / 4; /* CPUDenseGraphAnnealer.cpp:56, CUDADenseGraphAnnealer.cu:143 */
    d_J_.allocate((size_t)N_ * Np_);
    d_h_.allocate((size_t)Np_);
    HIP_CHECK(hipMemsetAsync(d_J_.ptr, 0, d_J_.bytes(), dev_->stream()));
    HIP_CHECK(hipMemsetAsync(d_h_.ptr, 0, d_h_.bytes(), dev_->stream()));
}

template <class real>
void DenseGraphAnnealer<real>::uploadJ(const real *J, const real *h, real c) {
    HIP_CHECK(hipMemcpy2DAsync(d_J_.ptr, (size_t)Np_ * sizeof(real), J, (size_t)N_ * sizeof(real),
                               (size_t)N_ * sizeof(real), N_, hipMemcpyHostToDevice, dev_->stream()));
    HIP_CHECK(hipMemcpyAsync(d_h_.ptr, h, (size_t)N_ * sizeof(real), hipMemcpyHostToDevice, dev_->stream()));
    prepareFieldForm();
    fragValid_ = false;
    planesValid_ = false;
    c_ = c;
    h_.assign(h, h + N_);
    dev_->synchronize();
    fieldsDirty_ = true;
}

template <class real>
void DenseGraphAnnealer<real>::prepareFieldForm() {
    if (fieldFormat_ == fieldF32) {
        d_Jsplit_.release();
        d_J8_.release();
        d_scale_.release();
        d_rowExp_.release();
        return;
    }
    const float *J = (const float *)d_J_.ptr;
    d_Jsplit_.allocate((size_t)N_ * Np_);
    if (fieldFormat_ == fieldBF16x2) {
        splitBF16(J, (size_t)N_ * Np_, d_Jsplit_.ptr, dev_->stream());
        d_J8_.release();
        d_scale_.release();
        d_rowExp_.release();
    } else {
        ld8_ = fp8RowStride(Np_);
        d_J8_.allocate((size_t)N_ * ld8_);
        d_scale_.allocate((size_t)Np_);
        d_rowExp_.allocate((size_t)N_);
        quantizeFP8(J, N_, Np_, ld8_, d_rowExp_.ptr, d_scale_.ptr, d_J8_.ptr, d_Jsplit_.ptr, dev_->stream());
    }
}

template <class real>
void DenseGraphAnnealer<real>::hamiltonianOf(const real *W, int N, OptimizeMethod om, bool check,
                                             std::vector<real> &h, std::vector<real> &J, real &c) {
    /* DGFuncs::calculateHamiltonian (SharedFormulas.cpp:36-55): h = -1/2
     * colsum W, J = -1/4 W, c = sum J + trace J, diag J = 0; maximize negates
     * h, J, c (CPUDenseGraphAnnealer.cpp:52-69).  Sums are accumulated in
     * double and rounded once. */
    SQ_THROW_IF(N <= 0, "W must be a non-empty square matrix.");
    if (check)
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < i; ++j)
                SQ_THROW_IF(W[(size_t)i * N + j] != W[(size_t)j * N + i], "W is not symmetric.");
    const real sgn = (om == optMaximize) ? real(-1.) : real(1.);
    h.assign(N, real(0.));
    J.assign((size_t)N * N, real(0.));
    double jsum = 0., dsum = 0.;
    for (int j = 0; j < N; ++j) {
        double s = 0.;
        for (int i = 0; i < N; ++i)
            s += (double)W[(size_t)i * N + j];
        h[j] = sgn * (real)(-0.5 * s);
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            real v = real(-0.25) * W[(size_t)i * N + j];
            jsum += (double)v;
            if (i == j)
                dsum += (double)v;
            J[(size_t)i * N + j] = (i == j) ? real(0.) : sgn * v;
        }
    c = sgn * (real)(jsum + dsum);
}

template <class real>
void DenseGraphAnnealer<real>::setProblem(const real *h, const real *J, real c, int N, OptimizeMethod om) {
    clearState(solProblemSet);
    allocateProblem(N);
    om_ = om;
    uploadJ(J, h, c);
    setState(solProblemSet);
}

template <class real>
void DenseGraphAnnealer<real>::setQUBO(const real *W, int N, OptimizeMethod om) {
    /* CPUDenseGraphAnnealer.cpp:52-69 */
    std::vector<real> h, J;
    real c;
    hamiltonianOf(W, N, om, true, h, J, c);
    setProblem(h.data(), J.data(), c, N, om);
}

template <class real>
void DenseGraphAnnealer<real>::setHamiltonian(const real *h, const real *J, real c, int N) {
    /* CPUDenseGraphAnnealer.cpp:71-84 (CUDA also requires symmetric J) */
    SQ_THROW_IF(N <= 0, "J must be a non-empty square matrix.");
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < i; ++j)
            SQ_THROW_IF(J[(size_t)i * N + j] != J[(size_t)j * N + i], "J is not symmetric.");
    clearState(solProblemSet);
    allocateProblem(N);
    om_ = optMinimize;
    /* J is kept as given, diagonal included: the reference's per-attempt dot
     * includes J[x][x] q[x] (CPUDenseGraphAnnealer.cpp:257), and so do the
     * local fields and their row updates here. */
    uploadJ(J, h, c);
    setState(solProblemSet);
}

template <class real>
void DenseGraphAnnealer<real>::getHamiltonian(real *h, real *J, real *c) const {
    throwErrorIfProblemNotSet();
    dev_->makeCurrent();
    memcpy(h, h_.data(), (size_t)N_ * sizeof(real));
    HIP_CHECK(hipMemcpy2DAsync(J, (size_t)N_ * sizeof(real), d_J_.ptr, (size_t)Np_ * sizeof(real),
                               (size_t)N_ * sizeof(real), N_, hipMemcpyDeviceToHost, dev_->stream()));
    dev_->synchronize();
    *c = c_;
}

template <class real>
void DenseGraphAnnealer<real>::prepare() {
    /* CUDADenseGraphAnnealer.cu:275-336 / CPUDenseGraphAnnealer.cpp:164-195 */
    throwErrorIfProblemNotSet();
    SQ_THROW_IF(m_ <= 0, "# trotters must be a positive integer.");
    ensureDevice();
    if (!isRandSeedGiven())
        seed((unsigned long long)time(NULL));
    setState(solRandSeedGiven);
    if (m_ == 1 && algo_ != algoSANaive) {
        algo_ = algoSANaive; /* selectDefaultSAAlgorithm, Solver.cpp:170-176 */
        logf("Selecting %s as the default SA algorithm.", algorithmToString(algo_));
    }
    const size_t mN = (size_t)m_ * N_;
    d_Q_.allocate((size_t)m_ * Nw_);
    d_F_.allocate((size_t)m_ * Np_);
    d_X_.allocate(mN);
    d_U_.allocate(mN);
    d_D_.allocate(mN);
    HIP_CHECK(hipMemsetAsync(d_Q_.ptr, 0, d_Q_.bytes(), dev_->stream()));
    HIP_CHECK(hipMemsetAsync(d_D_.ptr, 0, d_D_.bytes(), dev_->stream()));
    tag_ = 0;
    d_E_.allocate((size_t)m_);
    ensureEnergyPlanes();
    d_spinDot_.allocate(1);
    d_abort_.allocate(1);
    d_counters_.allocate(16); /* [0..2] public, [3..15] optional per-phase cycles */
    HIP_CHECK(hipMemsetAsync(d_abort_.ptr, 0, sizeof(int), dev_->stream()));
    HIP_CHECK(hipMemsetAsync(d_counters_.ptr, 0, 16 * sizeof(unsigned long long), dev_->stream()));
    h_abort_.allocate(1, hipHostMallocCoherent | hipHostMallocMapped);
    h_abort_.ptr[0] = 0;
    HIP_CHECK(hipHostGetDevicePointer((void **)&d_abortHost_, h_abort_.ptr, 0));
    abortPending_ = false;
    if (rngMT_) {
        /* the generators continue across prepare() calls, as the CPU
         * annealer's do; they are (re)seeded by seed() and on first use */
        if (!d_mt_.ptr || mtWorkersSeeded_ != nWorkers_)
            seedWorkerGenerators();
        /* the naive step draws one word more per attempt (its slice); the
         * worker streams of the OpenMP order each hold a step's draws */
        const int perAttempt = sizeof(real) == 4 ? 3 : 4;
        size_t words = mN * perAttempt;
        if (nWorkers_ > 1)
            words = std::max<size_t>(words, (size_t)nWorkers_ * mtWorkerStride(N_, m_, nWorkers_, perAttempt - 1));
        d_words_.allocate(words);
    }
    E_.assign((size_t)m_, real(0.));
    q_.assign(mN, 0);
    dev_->synchronize();
    fieldsDirty_ = true;
    setState(solPrepared);
}

template <class real>
void DenseGraphAnnealer<real>::randomizeSpin() {
    /* CPUDenseGraphAnnealer.cpp:151-161: q = randInt(2) ? +1 : -1, row-major */
    throwErrorIfNotPrepared();
    ensureDevice();
    if (rngMT_) {
        mtGenerate(d_mt_.ptr, d_words_.ptr, (size_t)m_ * N_, dev_->stream());
        spinsFromWords(d_words_.ptr, N_, N_, m_, Nw_, d_Q_.ptr, dev_->stream());
    } else {
        spinsFromPhilox(seed_, sweepIdx_++, N_, m_, Nw_, d_Q_.ptr, dev_->stream());
    }
    fieldsDirty_ = true;
    sweepsSinceSpinsSet_ = 0;
    setState(solQSet);
}

template <class real>
void DenseGraphAnnealer<real>::set_q(const int8_t *q, int N) {
    /* CUDADenseGraphAnnealer.cu:200-211: one q broadcast to every slice */
    throwErrorIfNotPrepared();
    SQ_THROW_IF(N != N_, "Dimension of q, %d, should be equal to N, %d.", N, N_);
    ensureDevice();
    std::vector<int8_t> all((size_t)m_ * N_);
    for (int y = 0; y < m_; ++y)
        memcpy(&all[(size_t)y * N_], q, (size_t)N_);
    d_q8_.allocate(all.size());
    HIP_CHECK(hipMemcpyAsync(d_q8_.ptr, all.data(), all.size(), hipMemcpyHostToDevice, dev_->stream()));
    packSpins(d_q8_.ptr, N_, m_, Nw_, d_Q_.ptr, dev_->stream());
    dev_->synchronize();
    fieldsDirty_ = true;
    sweepsSinceSpinsSet_ = 0;
    setState(solQSet);
}

template <class real>
void DenseGraphAnnealer<real>::set_qset(const int8_t *q, int m, int N) {
    /* CUDADenseGraphAnnealer.cu:213-230: m = number of vectors, prepare(), copy */
    throwErrorIfProblemNotSet();
    SQ_THROW_IF(N != N_, "Dimension of q, %d, should be equal to N, %d.", N, N_);
    SQ_THROW_IF(m <= 0, "empty q set");
    m_ = m;
    prepare();
    d_q8_.allocate((size_t)m_ * N_);
    HIP_CHECK(hipMemcpyAsync(d_q8_.ptr, q, (size_t)m_ * N_, hipMemcpyHostToDevice, dev_->stream()));
    packSpins(d_q8_.ptr, N_, m_, Nw_, d_Q_.ptr, dev_->stream());
    dev_->synchronize();
    fieldsDirty_ = true;
    sweepsSinceSpinsSet_ = 0;
    setState(solQSet);
}

template <class real>
void DenseGraphAnnealer<real>::checkAbort(bool synchronous) {
    /* The sweep writes the abort flag into pinned host memory.  Without a
     * synchronisation (an annealOneStep call) a flag already written is
     * reported early; synchronous callers have synchronised the stream first,
     * so every finished sweep is seen. */
    if (!abortPending_)
        return;
    const int flag = __atomic_load_n(h_abort_.ptr, __ATOMIC_ACQUIRE);
    if (synchronous)
        abortPending_ = false;
    if (flag) {
        dev_->synchronize();
        abortPending_ = false;
        h_abort_.ptr[0] = 0;
        HIP_CHECK(hipMemsetAsync(d_abort_.ptr, 0, sizeof(int), dev_->stream()));
        dev_->synchronize();
        clearState(solQSet);
        SQ_THROW("sweep aborted: a slice waited past the spin limit on a neighbour's decision; spins are invalid");
    }
}

template <class real>
void DenseGraphAnnealer<real>::refreshFields(bool always) {
    if (plainSweep_ && !always) {
        fieldsDirty_ = false;
        return;
    }
    /* F = Q . J in the field format: MFMA for float solvers, vector FMA for
     * double (and on request, as the cross-check of the MFMA kernels) */
    const bool mfma = sizeof(real) == 4 && !refreshVector_;
    if (mfma)
        ensureFragments();
    if (fieldFormat_ == fieldFP8) {
        /* the vector cross-check runs on J' as floats */
        if (mfma)
            fieldRefreshFP8(d_frag8_.ptr, d_frag8_.n, d_scale_.ptr, d_Q_.ptr, Np_, Np_, m_, Nw_, Np_,
                            (float *)d_F_.ptr, dev_->stream());
        else
            fieldRefresh<float>(d_Jsplit_.ptr, N_, Np_, Np_, d_Q_.ptr, m_, Nw_, Np_, (float *)d_F_.ptr, dev_->stream());
    } else if (fieldFormat_ == fieldBF16x2) {
        /* the vector cross-check runs on Jsplit = float(hi) + float(lo), the
         * same numbers as the two planes */
        if (mfma)
            fieldRefreshBF16x2(d_frag16_.ptr, d_frag16_.n, d_Q_.ptr, Np_, Np_, m_, Nw_, Np_, (float *)d_F_.ptr,
                               dev_->stream());
        else
            fieldRefresh<float>(d_Jsplit_.ptr, N_, Np_, Np_, d_Q_.ptr, m_, Nw_, Np_, (float *)d_F_.ptr, dev_->stream());
    } else if (mfma) {
        fieldRefreshF32MFMA(d_frag32_.ptr, d_frag32_.n, d_Q_.ptr, Np_, Np_, m_, Nw_, Np_, (float *)d_F_.ptr,
                            dev_->stream());
    } else {
        fieldRefresh<real>(d_J_.ptr, N_, Np_, Np_, d_Q_.ptr, m_, Nw_, Np_, d_F_.ptr, dev_->stream());
    }
    fieldsDirty_ = false;
    stepsSinceRefresh_ = 0;
}

template <class real>
void DenseGraphAnnealer<real>::generateAttempts() {
    const int order = (algo_ == algoSANaive) ? orderSliceMajor : orderColoring;
    if (rngMT_ && nWorkers_ > 1 && algo_ == algoColoring) {
        workerStreamsToAttempts<real>(d_mt_.ptr, d_words_.ptr, N_, m_, nWorkers_, d_X_.ptr, d_U_.ptr, dev_->stream());
    } else if (rngMT_) {
        const int perAttempt = sizeof(real) == 4 ? 2 : 3;
        mtGenerate(d_mt_.ptr, d_words_.ptr, (size_t)m_ * N_ * perAttempt, dev_->stream());
        streamToAttempts<real>(d_words_.ptr, N_, m_, order, d_X_.ptr, d_U_.ptr, dev_->stream());
    } else {
        philoxAttempts<real>(seed_, sweepIdx_++, N_, m_, d_X_.ptr, d_U_.ptr, dev_->stream());
    }
}

template <class real>
bool DenseGraphAnnealer<real>::refreshDue() {
    if (!fieldsDirty_ && refreshInterval_ > 0 && stepsSinceRefresh_ >= refreshInterval_)
        fieldsDirty_ = true;
    return fieldsDirty_ && !plainSweep_;
}

template <class real>
void DenseGraphAnnealer<real>::reserveTags(int n) {
    /* decision words carry the step's tag; before the tags wrap, the words of
     * every past step are cleared so none can match a new tag */
    if ((unsigned long long)tag_ + (unsigned long long)n >= 0x7fffffffull) {
        tag_ = 0;
        HIP_CHECK(hipMemsetAsync(d_D_.ptr, 0, d_D_.bytes(), dev_->stream()));
    }
}

template <class real>
SweepArgs<real> DenseGraphAnnealer<real>::makeSweepArgs(real G, real beta) {
    SweepArgs<real> a{};
    a.N = N_;
    a.Np = Np_;
    a.m = m_;
    a.Nw = Nw_;
    a.sa = (algo_ == algoSANaive) ? 1 : 0;
    /* host scalars in the reference's expression order, in real (:283-284, :358) */
    a.twoDivM = real(2.) / real(m_);
    a.coef = std::log(std::tanh(G * beta / real(m_))) / beta;
    a.beta = beta;
    a.invKT = real(1.) / G;
    a.J = (fieldFormat_ != fieldF32 && !plainSweep_) ? reinterpret_cast<const real *>(d_Jsplit_.ptr) : d_J_.ptr;
    a.J8 = d_J8_.ptr;
    a.rowScale = d_scale_.ptr;
    a.ld8 = ld8_;
    a.h = d_h_.ptr;
    a.F = d_F_.ptr;
    a.Qbits = d_Q_.ptr;
    a.X = d_X_.ptr;
    a.U = d_U_.ptr;
    a.D = d_D_.ptr;
    reserveTags(1);
    a.tag = ++tag_;
    a.abortFlag = d_abort_.ptr;
    a.abortHost = d_abortHost_;
    a.counters = d_counters_.ptr;
    a.spinLimit = spinLimit_;
    return a;
}

template <class real>
void DenseGraphAnnealer<real>::annealOneStepNaive(real G, real beta) {
    /* CPUDenseGraphAnnealer.cpp:268-280 on one workgroup (dense_naive.hip).
     * The step keeps local fields like the coloring sweep; the plain sweep
     * does not maintain them, so after plain steps they are recomputed. */
    if (plainSweep_)
        fieldsDirty_ = true;
    if (!fieldsDirty_ && refreshInterval_ > 0 && stepsSinceRefresh_ >= refreshInterval_)
        fieldsDirty_ = true;
    if (fieldsDirty_)
        refreshFields(true);
    NaiveArgs<real> a{};
    a.N = N_;
    a.Np = Np_;
    a.m = m_;
    a.Nw = Nw_;
    /* host scalars in the reference's expression order, in real (:272-273) */
    a.twoDivM = real(2.) / real(m_);
    a.coef = std::log(std::tanh(G * beta / real(m_))) / beta;
    a.beta = beta;
    a.J = (fieldFormat_ != fieldF32) ? reinterpret_cast<const real *>(d_Jsplit_.ptr) : d_J_.ptr;
    a.h = d_h_.ptr;
    a.F = d_F_.ptr;
    a.Qbits = d_Q_.ptr;
    a.counters = d_counters_.ptr;
    if (rngMT_) {
        const int words = sizeof(real) == 4 ? 3 : 4;
        mtGenerate(d_mt_.ptr, d_words_.ptr, (size_t)m_ * N_ * words, dev_->stream());
        a.words = d_words_.ptr;
    } else {
        const unsigned long long s = sweepIdx_++;
        a.words = nullptr;
        a.k0 = (uint32_t)seed_;
        a.k1 = (uint32_t)(seed_ >> 32);
        a.s0 = (uint32_t)s;
        a.s1 = (uint32_t)(s >> 32);
    }
    naiveSweep<real>(a, dev_->stream());
    ++stepsSinceRefresh_;
    clearState(solSolutionAvailable);
}

template <class real>
void DenseGraphAnnealer<real>::stepDone() {
    abortPending_ = true; /* the sweep reports an abort in h_abort_ itself */
    ++stepsSinceRefresh_;
    clearState(solSolutionAvailable);
}

template <class real>
void DenseGraphAnnealer<real>::annealOneStep(real G, real beta) {
    /* CPUDenseGraphAnnealer.cpp:281-301 (coloring) and :354-371 (SA) */
    throwErrorIfQNotSet();
    ensureDevice();
    checkAbort(false);
    if (algo_ == algoNaive) {
        annealOneStepNaive(G, beta);
        if (sysTraceCap_ > 0)
            traceSystemE(G, beta);
        return;
    }
    PhaseEvents pe;
    if (refreshDue()) {
        if (profile_ >= 2)
            phaseBegin(0, pe);
        refreshFields();
        if (profile_ >= 2)
            phaseEnd(pe);
    }
    if (profile_ >= 2)
        phaseBegin(1, pe);
    generateAttempts();
    if (profile_ >= 2)
        phaseEnd(pe);

    SweepArgs<real> a = makeSweepArgs(G, beta);
    int S = slicesPerBlock_;
    if (S <= 0) {
        /* one slice per workgroup while the CUs last (a CU to each wave is
         * faster than sharing one), more when m exceeds the CU count, as LDS
         * allows */
        S = 1;
        while (S < 4 && divru(m_, S) > dev_->numCUs())
            S <<= 1;
        while (S > 1 && denseSweepLdsBytes(Np_, S, sizeof(real), plainSweep_, fieldFormat_, ld8_) >
                            dev_->maxLdsPerBlock())
            S >>= 1;
    }
    SQ_THROW_IF(denseSweepLdsBytes(Np_, S, sizeof(real), plainSweep_, fieldFormat_, ld8_) > dev_->maxLdsPerBlock(),
                "N=%d too large for the LDS-resident sweep (%zu bytes needed per block)", N_,
                denseSweepLdsBytes(Np_, S, sizeof(real), plainSweep_, fieldFormat_, ld8_));
    /* multi-flip batches where LDS allows (float local-field sweeps).  The
     * fp8 form's run the float kernel on the float copy of J' (a.J): its row
     * updates add the same values, -+2 J'[x][s], exactly. */
    const bool multi = multiFlipMode_ == kMultiFlipOn ||
                       (multiFlipMode_ == kMultiFlipAuto && sweepsSinceSpinsSet_ < kMultiFlipAutoSweeps);
    int sweepFormat = fieldFormat_;
    a.multiMinCand = multiFlipMin_;
    a.multiBufs = 0;
    if (multi) {
        const int f = fieldFormat_ == fieldFP8 ? fieldF32 : fieldFormat_;
        a.multiBufs = denseSweepMultiBufs(Np_, S, sizeof(real), plainSweep_, f, dev_->maxLdsPerBlock());
        if (a.multiBufs >= 2)
            sweepFormat = f;
    }
    /* the helper-wave sweep where it fits; the fp8 form's runs on the float
     * copy of J' as well (the same row values, -+2 J'[x][s]) */
    if (helperWave_ && a.multiBufs < 2) {
        const int f = sweepFormat == fieldFP8 ? fieldF32 : sweepFormat;
        if (denseSweepHelperFits(Np_, S, sizeof(real), plainSweep_, f, dev_->maxLdsPerBlock())) {
            a.helperWave = 1;
            sweepFormat = f;
        }
    }
    ++sweepsSinceSpinsSet_;
    if (profile_)
        phaseBegin(2, pe);
    denseSweep<real>(a, plainSweep_, sweepFormat, S, !coopLaunch_ && dev_->soleSweeper(), dev_->numCUs(),
                     dev_->stream());
    if (profile_)
        phaseEnd(pe);
    if (pending_.size() > 3000)
        collectPhases();
    stepDone();
    if (sysTraceCap_ > 0)
        traceSystemE(G, beta);
}

template <class real>
bool DenseGraphAnnealer<real>::batchable(std::string *why) const {
    const char *r = nullptr;
    if (sizeof(real) != 4)
        r = "float32 solvers only";
    else if (rngMT_)
        r = "rng=philox only";
    else if (plainSweep_)
        r = "sweep=local only";
    else if (refreshVector_)
        r = "refresh_engine=mfma only";
    if (r && why)
        *why = r;
    return r == nullptr;
}

template <class real>
void DenseGraphAnnealer<real>::ensureFragments() {
    if (fragValid_ || sizeof(real) != 4)
        return;
    /* J in MFMA fragment order, built once per problem and form */
    const float *J = (const float *)d_J_.ptr;
    d_frag16_.release();
    d_frag32_.release();
    d_frag8_.release();
    if (fieldFormat_ == fieldBF16x2) {
        d_frag16_.allocate(fieldFragElemsBF16x2(Np_, Np_));
        buildFragBF16x2(J, N_, Np_, Np_, Np_, d_frag16_.ptr, d_frag16_.n, dev_->stream());
    } else if (fieldFormat_ == fieldFP8) {
        d_frag8_.allocate(fieldFragElemsFP8(Np_, Np_));
        buildFragFP8(d_J8_.ptr, N_, ld8_, Np_, Np_, d_frag8_.ptr, d_frag8_.n, dev_->stream());
    } else {
        d_frag32_.allocate(fieldFragElemsF32(Np_, Np_));
        buildFragF32(J, N_, Np_, Np_, Np_, d_frag32_.ptr, d_frag32_.n, dev_->stream());
    }
    fragValid_ = true;
}

template <class real>
FieldGemmJob DenseGraphAnnealer<real>::takeFieldGemmJob() {
    ensureFragments();
    FieldGemmJob j{};
    j.frag = fieldFormat_ == fieldBF16x2 ? (const void *)d_frag16_.ptr
             : fieldFormat_ == fieldFP8  ? (const void *)d_frag8_.ptr
                                         : (const void *)d_frag32_.ptr;
    j.colScale = d_scale_.ptr;
    j.Qbits = d_Q_.ptr;
    j.F = (float *)d_F_.ptr;
    j.nCols = Np_;
    j.K = Np_;
    j.m = m_;
    j.Nw = Nw_;
    j.ldF = Np_;
    checkFieldGemmJob(j, fieldFormat_);
    fieldsDirty_ = false;
    stepsSinceRefresh_ = 0;
    return j;
}

template <class real>
AttemptStreamJob DenseGraphAnnealer<real>::takeAttemptStreamJob() {
    AttemptStreamJob j{};
    j.seed = seed_;
    j.sweepIdx = sweepIdx_++;
    j.N = N_;
    j.m = m_;
    j.X = d_X_.ptr;
    j.U = d_U_.ptr;
    return j;
}

template <class real>
void DenseGraphAnnealer<real>::ensureEnergyPlanes() {
    /* the planes and the partial sums of the energy kernels for this J and m
     * (synchronises the stream when the planes are built) */
    if (!planesValid_) {
        d_planeScan_.allocate(3);
        planes_ = energyPlaneLayout<real>(d_J_.ptr, Np_, 1, N_, N_, d_planeScan_.ptr, dev_->stream());
        if (planes_.finite) {
            d_planes_.allocate(energyPlaneFragCount(planes_));
            energyPlaneBuild<real>(d_J_.ptr, Np_, 1, planes_, d_planes_.ptr, dev_->stream());
        } else {
            d_planes_.release();
        }
        planesValid_ = true;
    }
    size_t parts = (size_t)m_ * (Np_ / 256); /* the FP64 vector kernel's */
    if (planes_.finite)
        parts = std::max(parts, energyPlanePartialCount(planes_, m_));
    d_partial_.allocate(parts);
}

template <class real>
void DenseGraphAnnealer<real>::computeEnergies() {
    /* exact integer planes on the MFMA units; a J holding NaN or infinity
     * takes the FP64 vector kernel */
    ensureEnergyPlanes();
    if (planes_.finite)
        energyPlanes<real>(planes_, d_planes_.ptr, d_h_.ptr, nullptr, (double)c_, d_Q_.ptr, Nw_, d_Q_.ptr, Nw_, m_,
                           d_partial_.ptr, d_E_.ptr, dev_->stream());
    else
        energyFP64<real>(d_J_.ptr, d_h_.ptr, (double)c_, d_Q_.ptr, N_, Np_, m_, Nw_, d_partial_.ptr, d_E_.ptr,
                         dev_->stream());
}

template <class real>
void DenseGraphAnnealer<real>::calculate_E() {
    /* CUDADenseGraphAnnealer.cu:254-271; energies evaluated exactly (integer
     * planes) or to FP64 accuracy */
    throwErrorIfQNotSet();
    ensureDevice();
    computeEnergies();
    std::vector<double> e((size_t)m_);
    HIP_CHECK(hipMemcpyAsync(e.data(), d_E_.ptr, (size_t)m_ * sizeof(double), hipMemcpyDeviceToHost, dev_->stream()));
    dev_->synchronize();
    checkAbort(true);
    const real sgn = (om_ == optMaximize) ? real(-1.) : real(1.);
    lastERoundErr_ = 0.;
    E_.resize((size_t)m_);
    Efp64_.resize((size_t)m_);
    for (int y = 0; y < m_; ++y) {
        E_[y] = sgn * (real)e[y];
        Efp64_[y] = (double)sgn * e[y];
        double err = std::fabs((double)(real)e[y] - e[y]);
        if (err > lastERoundErr_)
            lastERoundErr_ = err;
    }
    setState(solEAvailable);
}

template <class real>
double DenseGraphAnnealer<real>::formEnergyError() {
    throwErrorIfQNotSet();
    ensureDevice();
    if constexpr (sizeof(real) != 4) {
        return 0.;
    } else {
        if (fieldFormat_ == fieldF32)
            return 0.;
        std::vector<double> e((size_t)m_), ef((size_t)m_);
        computeEnergies();
        HIP_CHECK(hipMemcpyAsync(e.data(), d_E_.ptr, (size_t)m_ * sizeof(double), hipMemcpyDeviceToHost,
                                 dev_->stream()));
        dev_->synchronize();
        /* the form's J' as digit planes of its own, for this call only */
        const EnergyPlaneLayout Lf =
            energyPlaneLayout<float>(d_Jsplit_.ptr, Np_, 1, N_, N_, d_planeScan_.ptr, dev_->stream());
        if (Lf.finite && planes_.finite) {
            DevBuf<uint4> frag;
            frag.allocate(energyPlaneFragCount(Lf));
            energyPlaneBuild<float>(d_Jsplit_.ptr, Np_, 1, Lf, frag.ptr, dev_->stream());
            if (energyPlanePartialCount(Lf, m_) > d_partial_.n) {
                dev_->synchronize(); /* d_partial_ is reallocated */
                d_partial_.allocate(energyPlanePartialCount(Lf, m_));
            }
            energyPlanes<float>(Lf, frag.ptr, d_h_.ptr, nullptr, (double)c_, d_Q_.ptr, Nw_, d_Q_.ptr, Nw_, m_,
                                d_partial_.ptr, d_E_.ptr, dev_->stream());
            HIP_CHECK(hipMemcpyAsync(ef.data(), d_E_.ptr, (size_t)m_ * sizeof(double), hipMemcpyDeviceToHost,
                                     dev_->stream()));
            dev_->synchronize(); /* before frag is freed */
        } else {
            energyFP64<real>(d_J_.ptr, d_h_.ptr, (double)c_, d_Q_.ptr, N_, Np_, m_, Nw_, d_partial_.ptr, d_E_.ptr,
                             dev_->stream());
            HIP_CHECK(hipMemcpyAsync(e.data(), d_E_.ptr, (size_t)m_ * sizeof(double), hipMemcpyDeviceToHost,
                                     dev_->stream()));
            dev_->synchronize();
            energyFP64<real>(d_Jsplit_.ptr, d_h_.ptr, (double)c_, d_Q_.ptr, N_, Np_, m_, Nw_, d_partial_.ptr,
                             d_E_.ptr, dev_->stream());
            HIP_CHECK(hipMemcpyAsync(ef.data(), d_E_.ptr, (size_t)m_ * sizeof(double), hipMemcpyDeviceToHost,
                                     dev_->stream()));
        }
        dev_->synchronize();
        checkAbort(true);
        double err = 0.;
        for (int y = 0; y < m_; ++y)
            err = std::max(err, std::fabs(ef[y] - e[y]));
        return err;
    }
}

template <class real>
void DenseGraphAnnealer<real>::makeSolution() {
    /* CPUDenseGraphAnnealer.cpp:197-227: syncBits then calculate_E */
    throwErrorIfQNotSet();
    ensureDevice();
    d_q8_.allocate((size_t)m_ * N_);
    unpackSpins(d_Q_.ptr, N_, m_, Nw_, d_q8_.ptr, dev_->stream());
    q_.resize((size_t)m_ * N_);
    HIP_CHECK(hipMemcpyAsync(q_.data(), d_q8_.ptr, q_.size(), hipMemcpyDeviceToHost, dev_->stream()));
    dev_->synchronize();
    checkAbort(true);
    setState(solSolutionAvailable);
    calculate_E();
}

template <class real>
void DenseGraphAnnealer<real>::get_q(int8_t *q) {
    if (!isSolutionAvailable())
        makeSolution();
    memcpy(q, q_.data(), q_.size());
}

template <class real>
void DenseGraphAnnealer<real>::get_x(int8_t *x) {
    /* x = (q + 1) / 2, Common.cpp:133-145 */
    if (!isSolutionAvailable())
        makeSolution();
    for (size_t i = 0; i < q_.size(); ++i)
        x[i] = (int8_t)((q_[i] + 1) / 2);
}

template <class real>
void DenseGraphAnnealer<real>::get_E(real *E) {
    if (!isEAvailable())
        calculate_E();
    memcpy(E, E_.data(), (size_t)m_ * sizeof(real));
}

template <class real>
real DenseGraphAnnealer<real>::getSystemE(real G, real beta) {
    /* CPUDenseGraphAnnealer.cpp:230-247 */
    calculate_E();
    real E = real(0.);
    for (int y = 0; y < m_; ++y)
        E += E_[y];
    E = E / real(m_);
    if (isSQAAlgorithm(algo_)) {
        double dot = 0.;
        spinDotSum(d_Q_.ptr, N_, m_, Nw_, d_spinDot_.ptr, dev_->stream());
        HIP_CHECK(hipMemcpyAsync(&dot, d_spinDot_.ptr, sizeof(double), hipMemcpyDeviceToHost, dev_->stream()));
        dev_->synchronize();
        real spinDotSum_ = (real)dot;
        real coef = real(0.5) / beta * std::log(std::tanh(G * beta / real(m_)));
        E -= spinDotSum_ * coef;
    }
    if (om_ == optMaximize)
        E *= real(-1.);
    return E;
}

template <class real>
void DenseGraphAnnealer<real>::setSystemETrace(int capacity) {
    SQ_THROW_IF(capacity < 0, "trace capacity must be >= 0");
    ensureDevice();
    sysTraceCap_ = capacity;
    sysTraceLen_ = 0;
    sysTraceSteps_ = 0;
    if (capacity > 0)
        d_sysTrace_.allocate((size_t)capacity);
    else
        d_sysTrace_.release();
}

template <class real>
void DenseGraphAnnealer<real>::traceSystemE(real G, real beta) {
    /* what getSystemE(G, beta) returns for the state after this step
     * (CPUDenseGraphAnnealer.cpp:230-247), computed on the device into the
     * trace with no host synchronisation */
    ++sysTraceSteps_;
    if (sysTraceLen_ >= sysTraceCap_)
        return;
    computeEnergies();
    const int sqa = isSQAAlgorithm(algo_) ? 1 : 0;
    real coef = real(0.);
    if (sqa) {
        spinDotSum(d_Q_.ptr, N_, m_, Nw_, d_spinDot_.ptr, dev_->stream());
        coef = real(0.5) / beta * std::log(std::tanh(G * beta / real(m_)));
    }
    const real sgn = (om_ == optMaximize) ? real(-1.) : real(1.);
    systemEnergy<real>(d_E_.ptr, m_, sgn, d_spinDot_.ptr, sqa, coef, sgn, d_sysTrace_.ptr + sysTraceLen_,
                       dev_->stream());
    ++sysTraceLen_;
}

template <class real>
int DenseGraphAnnealer<real>::getSystemETrace(double *out, int cap, long long *steps) {
    ensureDevice();
    const int n = std::min(cap, sysTraceLen_);
    std::vector<real> v((size_t)std::max(n, 0));
    if (n > 0) {
        HIP_CHECK(hipMemcpyAsync(v.data(), d_sysTrace_.ptr, (size_t)n * sizeof(real), hipMemcpyDeviceToHost,
                                 dev_->stream()));
        dev_->synchronize();
    }
    for (int i = 0; i < n; ++i)
        out[i] = (double)v[i];
    if (steps)
        *steps = sysTraceSteps_;
    return n;
}

template <class real>
void DenseGraphAnnealer<real>::getFields(real *F) {
    throwErrorIfQNotSet();
    SQ_THROW_IF(plainSweep_, "the plain sweep keeps no local fields");
    ensureDevice();
    if (fieldsDirty_)
        refreshFields();
    HIP_CHECK(hipMemcpy2DAsync(F, (size_t)N_ * sizeof(real), d_F_.ptr, (size_t)Np_ * sizeof(real),
                               (size_t)N_ * sizeof(real), m_, hipMemcpyDeviceToHost, dev_->stream()));
    dev_->synchronize();
    checkAbort(true);
}

template <class real>
void DenseGraphAnnealer<real>::getCounters(unsigned long long out[3]) {
    out[0] = out[1] = out[2] = 0;
    if (!d_counters_.ptr)
        return;
    ensureDevice();
    HIP_CHECK(hipMemcpyAsync(out, d_counters_.ptr, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             dev_->stream()));
    dev_->synchronize();
}

template class DenseGraphAnnealer<float>;
template class DenseGraphAnnealer<double>;

}  // namespace sqaod_rocm
