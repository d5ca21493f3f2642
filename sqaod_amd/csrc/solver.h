/*
 * Host-side solver objects of the rocm device.  They mirror the reference's
 * abstract solver API (sqaodc/common/Solver.h:17-252: Solver / Annealer /
 * DenseGraphAnnealer / BipartiteGraphAnnealer / DenseGraphBFSearcher<real>)
 * and its state machine (sqaodc/common/Solver.cpp:32-112), with device-
 * resident state driven through the HIP kernels of this directory.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace sqaod_rocm {

/* sqaodc/common/Preference.h:8-17 */
enum Algorithm {
    algoUnknown = 0,
    algoDefault = 1,
    algoNaive = 2,
    algoColoring = 3,
    algoBruteForceSearch = 4,
    algoSADefault = 5,
    algoSANaive = 6,
    algoSAColoring = 7,
};
const char *algorithmToString(Algorithm a);
Algorithm algorithmFromString(const char *s);
bool isSQAAlgorithm(Algorithm a);

/* sqaodc/common/Solver.h:9-14 */
enum OptimizeMethod { optMinimize = 0, optMaximize = 1 };

/* Logging to stderr when SQAOD_VERBOSE is non-zero (defines.cpp:46-63). */
void logf(const char *fmt, ...);

/* A GPU and the stream its solvers enqueue on (cuda/Device.cpp role). */
class Device {
public:
    explicit Device(int devno);
    ~Device();
    int devno() const { return devno_; }
    hipStream_t stream() const { return stream_; }
    int numCUs() const { return numCUs_; }
    size_t maxLdsPerBlock() const { return maxLds_; }
    void makeCurrent() const;
    void synchronize() const;
    /* Registers this Device (its stream) as one that launches spin-waiting
     * sweeps and returns whether it is the only such Device of its GPU in this
     * process: then no sweep of another stream can run beside its sweeps. */
    bool soleSweeper();

private:
    int devno_;
    hipStream_t stream_;
    bool sweeper_ = false;
    int numCUs_ = 0;
    size_t maxLds_ = 65536;
};

/* The process-wide default device (device 0), created on first use, like
 * sqaod.cuda's active_device (sqaodpy/sqaod/cuda/device.py:26-37). */
Device &defaultDevice();

/* Solver state machine, Solver.cpp:32-112. */
class SolverState {
public:
    enum {
        solNone = 0,
        solProblemSet = 1,
        solPrepared = 2,
        solEAvailable = 4,
        solSolutionAvailable = 8,
        solRandSeedGiven = 16,
        solQSet = 32,
    };
    void setState(int s) {
        clearState(s);
        state_ |= s;
    }
    void clearState(int s);
    bool isRandSeedGiven() const { return state_ & solRandSeedGiven; }
    bool isProblemSet() const { return state_ & solProblemSet; }
    bool isPrepared() const { return state_ & solPrepared; }
    bool isQSet() const { return state_ & solQSet; }
    bool isEAvailable() const { return state_ & solEAvailable; }
    bool isSolutionAvailable() const { return state_ & solSolutionAvailable; }
    void throwErrorIfProblemNotSet() const;
    void throwErrorIfNotPrepared() const;
    void throwErrorIfQNotSet() const;

protected:
    int state_ = solNone;
};

using PrefList = std::vector<std::pair<std::string, std::string>>;

/* ------------------------------------------------------------------ */
/* Dense-graph annealer (CPUDenseGraphAnnealer / CUDADenseGraphAnnealer)   */
/* ------------------------------------------------------------------ */
template <class real>
class DenseGraphAnnealer : public SolverState {
public:
    DenseGraphAnnealer();
    ~DenseGraphAnnealer();

    void assignDevice(Device &dev);

    void seed(unsigned long long s);
    Algorithm selectAlgorithm(Algorithm algo);
    Algorithm getAlgorithm() const { return algo_; }
    void setPreference(const std::string &name, const std::string &value);
    PrefList getPreferences() const;

    void setQUBO(const real *W, int N, OptimizeMethod om);
    void setHamiltonian(const real *h, const real *J, real c, int N);
    /* setQUBO in two halves: the host Hamiltonian (h, J N x N with a zero
     * diagonal, c; a pure function, callable from any thread; check = test W
     * for symmetry) and its upload as the annealer's problem */
    static void hamiltonianOf(const real *W, int N, OptimizeMethod om, bool check, std::vector<real> &h,
                              std::vector<real> &J, real &c);
    void setProblem(const real *h, const real *J, real c, int N, OptimizeMethod om);
    void getHamiltonian(real *h, real *J, real *c) const;
    int getProblemSize() const { return N_; }
    int numTrotters() const { return m_; }

    void set_q(const int8_t *q, int N);
    void set_qset(const int8_t *q, int m, int N);
    void get_q(int8_t *q);             /* m x N */
    void get_x(int8_t *x);             /* m x N */
    void get_E(real *E);               /* m */
    void randomizeSpin();
    void prepare();
    void calculate_E();
    void makeSolution();
    real getSystemE(real G, real beta);
    void annealOneStep(real G, real beta);

    /* diagnostics (not in the reference API) */
    void getCounters(unsigned long long out[3]);
    /* With the "profile" preference on, HIP events bracket each phase of
     * annealOneStep on the solver's stream; this returns the summed device
     * times (ms) and launch counts of [field refresh, stream generation,
     * sweep] since the last call, and resets them. */
    void getPhaseTimes(double ms[3], unsigned long long counts[3]);
    /* The local fields F = J q_y (m x N, real) as the sweep holds them after
     * the last step (recomputed first when the spins were set since). */
    void getFields(real *F);
    /* Max |E_real - E_fp64| of the last calculate_E (the rounding of the
     * FP64 recompute to real). */
    double lastEnergyRoundingError() const { return lastERoundErr_; }
    /* The energies of the last calculate_E before their rounding to real. */
    const std::vector<double> &energiesFP64() const { return Efp64_; }
    /* Max over slices of |E_form(q_y) - E(q_y)| for the current spins, both
     * in FP64: E_form uses the J the field form anneals (bf16 hi + lo, or the
     * fp8 J'), E the J as given; 0 for the fp32/fp64 form. */
    double formEnergyError();
    /* System-energy trace (no reference counterpart): with capacity > 0,
     * every annealOneStep(G, beta) appends getSystemE(G, beta) of the state
     * after the step, computed on the device without a host synchronisation;
     * the first `capacity` steps are kept.  0 turns it off.  Reading copies
     * the kept values (one synchronisation) and the number of steps traced. */
    void setSystemETrace(int capacity);
    int getSystemETrace(double *out, int cap, long long *steps);

    /* Hooks of the batch engine (dense_batch.cpp), which runs many
     * instances' steps in shared launches with the same kernels and the same
     * arguments annealOneStep uses. */
    bool batchable(std::string *why) const; /* float, philox, local sweep, MFMA refresh */
    bool refreshDue();                       /* the refresh policy of annealOneStep */
    void reserveTags(int n);                 /* clears the decision words before the tag wraps */
    FieldGemmJob takeFieldGemmJob();         /* the step's field refresh, enqueued by the caller */
    AttemptStreamJob takeAttemptStreamJob(); /* the step's Philox streams (advances the sweep index) */
    SweepArgs<real> makeSweepArgs(real G, real beta);
    void stepDone();
    void checkStepAbort() { checkAbort(false); }
    int fieldFormat() const { return fieldFormat_; }
    int paddedN() const { return Np_; }
    int fp8RowStrideBytes() const { return ld8_; }
    Device *device() const { return dev_; }
    /* spins and fields, for snapshots around trial steps */
    uint32_t *spinWords() { return d_Q_.ptr; }
    real *fields() { return d_F_.ptr; }
    struct StepState {
        unsigned long long sweepIdx;
        int stepsSinceRefresh;
        bool fieldsDirty;
    };
    StepState stepState() const { return {sweepIdx_, stepsSinceRefresh_, fieldsDirty_}; }
    void setStepState(const StepState &s) {
        sweepIdx_ = s.sweepIdx;
        stepsSinceRefresh_ = s.stepsSinceRefresh;
        fieldsDirty_ = s.fieldsDirty;
    }

private:
    void ensureDevice();
    void checkAbort(bool synchronous);
    void allocateProblem(int N);
    void refreshFields(bool always = false); /* always: also when the plain sweep needs no fields */
    void annealOneStepNaive(real G, real beta);
    void seedWorkerGenerators(); /* d_mt_[w] = MT19937 seeded seed_ + 17 w, w < nWorkers_ */
    void traceSystemE(real G, real beta);
    void generateAttempts();
    void uploadJ(const real *J, const real *h, real c);
    void prepareFieldForm(); /* derives the form's J (bf16 split / fp8 quantization) from d_J_ */
    void ensureFragments();  /* J in MFMA fragment order for the field form */

    Device *dev_ = nullptr;
    int N_ = 0, Np_ = 0, Nw_ = 0, m_ = -1;
    Algorithm algo_ = algoColoring;
    OptimizeMethod om_ = optMinimize;
    unsigned long long seed_ = 0;
    unsigned long long sweepIdx_ = 0;
    uint32_t tag_ = 0;
    real c_ = 0;

    /* preferences beyond the reference's */
    bool rngMT_ = false;          /* "rng": "philox" (default) | "mt19937" (CPU-stream replay) */
    bool plainSweep_ = false;     /* "sweep": "local" (default) | "plain" */
    int fieldFormat_ = fieldF32;  /* "field_precision": "fp32" | "bf16x2" | "fp8" */
    int slicesPerBlock_ = 0;      /* "slices_per_block": 0 = automatic */
    bool coopLaunch_ = false;     /* "sweep_launch": cooperative always, or direct when safe (auto) */
    int refreshInterval_ = 1;     /* "refresh_interval": steps between F recomputes (0: only when spins were set) */
    bool refreshVector_ = false;  /* "refresh_engine": "mfma" (default) | "vector" */
    int spinLimit_ = 1 << 20;
    int nWorkers_ = 1;        /* CPU workers whose MT19937 streams the coloring step replays */
    /* multi-flip batches in the float local-field sweep: off, on, or (auto)
     * in the first kMultiFlipAutoSweeps sweeps after the spins were set;
     * with them, a batch with at least multiFlipMin_ accepting lanes makes
     * the next one resolve all its flips in order */
    int multiFlipMode_ = kMultiFlipAuto;
    int multiFlipMin_ = kMultiFlipDefaultMin;
    int sweepsSinceSpinsSet_ = 0;
    /* single-flip float local-field sweeps with one slice per block: a second
     * wave per slice applies the deferred J rows (where LDS allows) */
    bool helperWave_ = true;
    int sysTraceCap_ = 0, sysTraceLen_ = 0;
    long long sysTraceSteps_ = 0;
    int mtWorkersSeeded_ = 0; /* generators in d_mt_ */

    int profile_ = 0;             /* "profile": 1 = HIP events around each sweep, 2 = around every phase */
    struct PhaseEvents {
        int phase;
        hipEvent_t start, stop;
    };
    std::vector<PhaseEvents> pending_;
    std::vector<hipEvent_t> eventPool_;
    double phaseMs_[3] = {0, 0, 0};
    unsigned long long phaseCount_[3] = {0, 0, 0};
    hipEvent_t takeEvent();
    void phaseBegin(int phase, PhaseEvents &pe);
    void phaseEnd(PhaseEvents &pe);
    void collectPhases();

    bool fieldsDirty_ = true;
    bool fragValid_ = false; /* d_frag16_/d_frag32_ hold the current J in MFMA fragment order */
    int stepsSinceRefresh_ = 0;
    double lastERoundErr_ = 0.;
    std::vector<double> Efp64_;

    /* device state */
    DevBuf<real> d_J_, d_h_, d_F_, d_U_;
    DevBuf<float> d_Jsplit_;  /* the form's J as floats: bf16x2 float(hi) + float(lo), fp8 J' */
    DevBuf<uint16_t> d_frag16_; /* bf16x2 MFMA refresh: J's hi/lo planes in fragment order */
    DevBuf<uint16_t> d_frag32_; /* fp32 MFMA refresh: J's exact hi/mid/lo bf16 planes in fragment order */
    DevBuf<uint8_t> d_J8_;      /* fp8 form: e4m3 rows (N x ld8_), row i scaled by d_scale_[i] */
    DevBuf<uint8_t> d_frag8_;   /* fp8 MFMA refresh: d_J8_ in fragment order */
    DevBuf<float> d_scale_;     /* fp8 form: 2^a_i per row (Np, 1 past N) */
    DevBuf<int> d_rowExp_;
    int ld8_ = 0;
    DevBuf<uint32_t> d_Q_, d_X_, d_D_, d_words_;
    DevBuf<MTState> d_mt_;
    DevBuf<double> d_E_, d_partial_, d_spinDot_;
    /* energies: J as int8 digit planes for the integer MFMA, built once per J */
    EnergyPlaneLayout planes_;
    DevBuf<uint4> d_planes_;
    DevBuf<int> d_planeScan_;
    bool planesValid_ = false;
    void ensureEnergyPlanes();
    void computeEnergies(); /* E of every slice into d_E_, on the stream */
    DevBuf<real> d_sysTrace_;
    DevBuf<int> d_abort_;
    DevBuf<unsigned long long> d_counters_;
    DevBuf<int8_t> d_q8_;
    PinnedBuf<int> h_abort_;
    int *d_abortHost_ = nullptr; /* device address of h_abort_ */
    bool abortPending_ = false;

    /* host copies */
    std::vector<real> h_;     /* host h (for getHamiltonian) */
    std::vector<real> E_;
    std::vector<int8_t> q_;   /* m x N after makeSolution */
};

/* ------------------------------------------------------------------ */
/* Batch of dense annealers stepped together (no reference counterpart) */
/* ------------------------------------------------------------------ */
/* Many independent dense instances (float, any N and m) advance through
 * one schedule together: per step one grouped field-GEMM launch, one
 * grouped Philox launch and one persistent sweep launch over every
 * (instance, slice), optionally captured as one hipGraph for the whole
 * schedule.  Each instance is a DenseGraphAnnealer<float> and its results
 * are bit-identical to running it alone with the same seed and preferences. */
class GroupedDenseStepper {
public:
    /* (G, beta) of step k for instance i */
    using Schedule = std::function<std::pair<double, double>(int, int)>;
    struct Options {
        int fieldFormat = fieldF32;
        int wavesPerBlock = 1;
        bool graph = false;
        bool wait = true;            /* synchronise and time the steps */
        double blockFraction = 1.0;  /* share of the resident blocks the persistent sweep may use */
    };
    /* nSteps steps of the instances listed in `order` (items of the
     * persistent sweep in that order); the others are left untouched */
    void run(Device &dev, const std::vector<DenseGraphAnnealer<float> *> &inst, const std::vector<int> &order,
             int nSteps, const Schedule &sched, const Options &o);
    double lastMs() const { return lastMs_; }
    /* instances whose sweep aborted (spin limit) in the last run: run()
     * checks every instance and then throws once if this is not empty */
    const std::vector<int> &lastAborted() const { return lastAborted_; }

private:
    DevBuf<unsigned char> d_plan_;
    PinnedBuf<unsigned char> h_plan_;
    DevBuf<int> d_queue_;
    double lastMs_ = 0.;
    std::vector<int> lastAborted_;
};

class DenseBatch {
public:
    explicit DenseBatch(Device &dev);
    ~DenseBatch();
    /* batch-wide preferences: field_precision, refresh_interval, spin_limit,
     * algorithm (as the annealer's), waves_per_block (1, 2, 4; 0 = tuned),
     * order ("largest_first" | "given") */
    void setPreference(const std::string &name, const std::string &value);
    PrefList getPreferences() const;
    /* setQUBO + n_trotters + seed on a new instance; returns its index */
    int addQUBO(const float *W, int N, int m, unsigned long long seed, OptimizeMethod om);
    int size() const { return (int)inst_.size(); }
    DenseGraphAnnealer<float> &instance(int i);
    void prepare(); /* prepare + randomizeSpin of every instance */
    /* steps k = 0..nSteps-1 with (G[k], beta[k]) on every instance; graph:
     * capture the whole schedule as one hipGraph and launch it once */
    void anneal(const double *G, const double *beta, int nSteps, bool graph);
    /* times trialSteps steps of each (waves per block, order) candidate at
     * (G, beta) and keeps the fastest; the instances' spins, fields and
     * streams are restored afterwards, so results do not depend on it */
    void autotune(double G, double beta, int trialSteps);
    long long attemptsPerStep() const;
    int wavesPerBlock() const { return wavesPerBlock_; }
    bool largestFirst() const { return largestFirst_; }
    double lastAnnealMs() const { return lastMs_; }

private:
    void buildItems();
    void run(const double *G, const double *beta, int nSteps, bool graph, int wpb);

    Device *dev_;
    std::vector<DenseGraphAnnealer<float> *> inst_;
    std::vector<std::pair<std::string, std::string>> prefs_;
    int fieldFormat_ = fieldF32;
    int wavesPerBlock_ = 0; /* 0: not tuned yet (1 is used) */
    bool largestFirst_ = true;
    bool itemsValid_ = false;
    double lastMs_ = 0.;
    GroupedDenseStepper stepper_;
    std::vector<int> order_;
    int maxNp_ = 0, maxLd8_ = 0, maxM_ = 0;
};

/* ------------------------------------------------------------------ */
/* Dense-graph brute-force searcher (CPUDenseGraphBFSearcher /           */
/* CUDADenseGraphBFSearcher)                                             */
/* ------------------------------------------------------------------ */
template <class real>
class DenseGraphBFSearcher : public SolverState {
public:
    using real_type = real;
    DenseGraphBFSearcher() = default;
    ~DenseGraphBFSearcher() = default;

    void assignDevice(Device &dev);
    void setPreference(const std::string &name, const std::string &value);
    PrefList getPreferences() const;

    void setQUBO(const real *W, int N, OptimizeMethod om);
    int getProblemSize() const { return N_; }
    void prepare();
    /* Scans the next span of x (max(tile size, 2^26) candidates, clipped to
     * 2^N) on the device without a host round trip; returns true when the
     * whole space has been scanned.  *curXEnd receives the end of the span. */
    bool searchRange(unsigned long long *curXEnd);
    void search();
    void makeSolution();
    void calculate_E();
    int numSolutions();           /* rows of get_x (makes the solution if needed) */
    int numEnergies();            /* length of get_E (calculates E if needed; at least 1) */
    void get_x(int8_t *x);        /* numSolutions() x N */
    void get_E(real *E);          /* numEnergies() */

private:
    void ensureDevice();

    Device *dev_ = nullptr;
    int N_ = 0;
    OptimizeMethod om_ = optMinimize;
    int tileSize_ = 1024; /* CPUDenseGraphBFSearcher.cpp:16 */
    std::vector<double> W_; /* N x N, negated for maximize (:37-38) */
    BFPlan plan_{};
    unsigned long long x_ = 0, xMax_ = 0, span_ = 0;
    int maxBlocks_ = 0;
    real Emin_ = 0;
    std::vector<unsigned long long> xList_;
    std::vector<real> E_;

    DevBuf<double> d_W_, d_EL_;
    DevBuf<BFState<real>> d_state_;
    DevBuf<real> d_blockMin_;
    DevBuf<unsigned> d_blockCnt_;
    DevBuf<unsigned long long> d_offsets_, d_xList_;
};

/* ------------------------------------------------------------------ */
/* Bipartite-graph brute-force searcher (CPUBipartiteGraphBFSearcher /   */
/* CUDABipartiteGraphBFSearcher)                                         */
/* ------------------------------------------------------------------ */
template <class real>
class BipartiteGraphBFSearcher : public SolverState {
public:
    using real_type = real;
    void assignDevice(Device &dev);
    void setPreference(const std::string &name, const std::string &value);
    PrefList getPreferences() const;
    /* W is N1 x N0, N0, N1 <= 63 */
    void setQUBO(const real *b0, const real *b1, const real *W, int N0, int N1, OptimizeMethod om);
    void getProblemSize(int *N0, int *N1) const;
    void prepare();
    /* scans the next rectangle of pairs; true once every pair is scanned;
     * *curX0 / *curX1: the position reached (x0 tile start, x1) */
    bool searchRange(unsigned long long *curX0, unsigned long long *curX1);
    void search();
    void makeSolution();
    void calculate_E();
    int numSolutions();
    int numEnergies();
    void get_x(int8_t *x0, int8_t *x1); /* numSolutions() x N0, x N1 */
    void get_E(real *E);

private:
    void ensureDevice();
    unsigned long long visitKey(unsigned long long x0, unsigned long long x1) const;

    Device *dev_ = nullptr;
    int N0_ = 0, N1_ = 0;
    OptimizeMethod om_ = optMinimize;
    int tileSize0_ = 1024, tileSize1_ = 1024; /* CPUBipartiteGraphBFSearcher.cpp:16-17 */
    unsigned long long T0_ = 1024;
    BGBFArgs plan_{};
    unsigned long long hNext_ = 0, x1Next_ = 0, x1Chunk_ = 0, hChunk_ = 0;
    real Emin_ = 0;
    std::vector<std::pair<unsigned long long, unsigned long long>> list_; /* (x0, x1), visiting order */
    std::vector<real> E_;
    DevBuf<real> d_b0_, d_b1_, d_W_, d_blockMin_;
    DevBuf<unsigned long long> d_pairs_;
    DevBuf<int> d_counts_;
};

/* ------------------------------------------------------------------ */
/* Bipartite-graph annealer (CPUBipartiteGraphAnnealer /                 */
/* CUDABipartiteGraphAnnealer)                                           */
/* ------------------------------------------------------------------ */
template <class real>
class BipartiteGraphAnnealer : public SolverState {
public:
    using real_type = real;
    BipartiteGraphAnnealer() = default;
    ~BipartiteGraphAnnealer() = default;

    void assignDevice(Device &dev);
    void seed(unsigned long long s);
    Algorithm selectAlgorithm(Algorithm algo);
    Algorithm getAlgorithm() const { return algo_; }
    void setPreference(const std::string &name, const std::string &value);
    PrefList getPreferences() const;

    /* W is N1 x N0 (rows: the x1 side), b0[N0], b1[N1] */
    void setQUBO(const real *b0, const real *b1, const real *W, int N0, int N1, OptimizeMethod om);
    void setHamiltonian(const real *h0, const real *h1, const real *J, real c, int N0, int N1);
    /* setQUBO in two halves: the host Hamiltonian in double (a pure
     * function, callable from any thread) and its upload as the problem */
    static void hamiltonianOf(const real *b0, const real *b1, const real *W, int N0, int N1, OptimizeMethod om,
                              std::vector<double> &h0, std::vector<double> &h1, std::vector<double> &J, double &c);
    void setProblem(const std::vector<double> &h0, const std::vector<double> &h1, const std::vector<double> &J,
                    double c, int N0, int N1, OptimizeMethod om);
    void getHamiltonian(real *h0, real *h1, real *J, real *c) const;
    void getProblemSize(int *N0, int *N1) const;
    int numTrotters() const { return m_; }

    void set_q(const int8_t *q0, int N0, const int8_t *q1, int N1); /* one pair for every slice */
    void set_qset(const int8_t *q0, const int8_t *q1, int m, int N0, int N1);
    void get_q(int8_t *q0, int8_t *q1); /* m x N0, m x N1 */
    void get_x(int8_t *x0, int8_t *x1);
    void get_E(real *E);
    void randomizeSpin();
    void prepare();
    void calculate_E();
    /* The energies of the last calculate_E before their rounding to real. */
    const std::vector<double> &energiesFP64() const { return Efp64_; }
    void makeSolution();
    real getSystemE(real G, real beta);
    void annealOneStep(real G, real beta);

private:
    void ensureDevice();
    void uploadProblem(const std::vector<double> &h0, const std::vector<double> &h1, const std::vector<double> &J,
                       double c);
    /* dEmat of one half-step: side 1 anneals q1 against q0 (dE = Q0 J^T),
     * side 0 anneals q0 against q1 (dE = Q1 J) */
    void computeFields(int side);
    void halfStep(int side, real G, real beta, const uint32_t *words);

    Device *dev_ = nullptr;
    int N0_ = 0, N1_ = 0, N0p_ = 0, N1p_ = 0, m_ = -1;
    Algorithm algo_ = algoColoring;
    OptimizeMethod om_ = optMinimize;
    unsigned long long seed_ = 0;
    unsigned long long stepIdx_ = 0;
    real c_ = 0;

    /* preferences beyond the reference's */
    bool rngMT_ = false;         /* "rng": "philox" (default) | "mt19937" */
    int fieldFormat_ = fieldF32; /* "field_precision": "fp32" | "bf16x2" (float solvers) */
    bool refreshVector_ = false; /* "refresh_engine": "mfma" | "vector" */
    bool fragValid_ = false;
    bool mtValid_ = false;       /* d_mt_ holds the generator seeded by the last seed() */

    std::vector<real> h0_, h1_;
    std::vector<real> E_;
    std::vector<double> Efp64_;
    std::vector<int8_t> q0_, q1_;

    DevBuf<real> d_J_, d_JT_, d_h0_, d_h1_, d_dE_;
    DevBuf<uint16_t> d_fragJ16_, d_fragJT16_;
    DevBuf<uint16_t> d_fragJ32_, d_fragJT32_;
    DevBuf<uint32_t> d_Q0_, d_Q1_, d_words_;
    DevBuf<MTState> d_mt_;
    DevBuf<double> d_E_, d_partial_, d_spinDot_;
    DevBuf<int8_t> d_q8_;
    /* energies: J (rows for q1) as int8 digit planes for the integer MFMA */
    EnergyPlaneLayout planes_;
    DevBuf<uint4> d_planes_;
    DevBuf<int> d_planeScan_;
    bool planesValid_ = false;
    void ensureEnergyPlanes();
    void computeEnergies();
};

/* ------------------------------------------------------------------ */
/* Formulas (DenseGraphFormulas / BipartiteGraphFormulas)                */
/* ------------------------------------------------------------------ */
template <class real>
class Formulas {
public:
    using real_type = real;
    void assignDevice(Device &dev);
    /* x, q: nBatch rows of N int8; E: nBatch */
    void denseCalculateE(real *E, const real *W, int N, const int8_t *x, int nBatch);
    void denseHamiltonian(real *h, real *J, real *c, const real *W, int N);
    void denseCalculateEFromSpin(real *E, const real *h, const real *J, real c, const int8_t *q, int N, int nBatch);
    /* W, J: N1 x N0; x0: nBatch0 x N0, x1: nBatch1 x N1; E: nBatch0 (pairs)
     * or nBatch1 x nBatch0 (twoD) */
    void bipartiteCalculateE(real *E, const real *b0, const real *b1, const real *W, int N0, int N1,
                             const int8_t *x0, int nBatch0, const int8_t *x1, int nBatch1, bool twoD);
    void bipartiteHamiltonian(real *h0, real *h1, real *J, real *c, const real *b0, const real *b1, const real *W,
                              int N0, int N1);
    void bipartiteCalculateEFromSpin(real *E, const real *h0, const real *h1, const real *J, real c, int N0, int N1,
                                     const int8_t *q0, const int8_t *q1, int nBatch);

private:
    Device &device();
    void run(QuadFormArgs<real> a, const real *A, size_t nA, const real *lin0, size_t n0, const real *lin1,
             size_t n1, const int8_t *U, size_t nU, const int8_t *V, size_t nV, real *E, int nOut);
    Device *dev_ = nullptr;
};

/* ------------------------------------------------------------------ */
/* Serving: continuous batching of anneal jobs (no reference counterpart) */
/* ------------------------------------------------------------------ */
/* Jobs (dense or bipartite float32 problems, each with its own seed and
 * G/beta schedule) are submitted at any time and admitted in arrival order
 * while the active-job limit and the device-memory budget allow.  A tick
 * advances every active job by one step: the dense jobs through grouped
 * launches on the server's stream (one persistent sweep per size class),
 * the bipartite jobs on a pool of other streams at the same time.  A job
 * that has run its schedule is finished (makeSolution: FP64 energies, x),
 * its device state is freed and its results wait to be collected.  Each
 * job's results equal the same job run alone with the same seed. */
class Server {
public:
    enum Kind { kDense = 0, kBipartite = 1 };
    struct Result {
        int kind, N0, N1, m;
        std::vector<float> E;
        std::vector<int8_t> x0, x1; /* dense: x0 is m x N */
        double submitS, admitS, finishS; /* seconds on the server's clock */
        long long attempts;
        std::string error; /* non-empty: the job failed (admission, a step or its solution threw) */
    };
    explicit Server(Device &dev);
    ~Server();
    /* memory_budget_mb, max_active, dense_field_precision, bipartite_field_precision,
     * bipartite_streams, block_fraction, dense_spin_limit */
    void setPreference(const std::string &name, const std::string &value);
    int submitDense(const float *W, int N, int m, unsigned long long seed, const double *G, const double *beta,
                    int nSteps);
    int submitBipartite(const float *b0, const float *b1, const float *W, int N0, int N1, int m,
                        unsigned long long seed, const double *G, const double *beta, int nSteps);
    /* admit, step every active job once, finish completed jobs; returns the
     * number of jobs still active or pending */
    int tick();
    int numPending() const { return (int)pending_.size(); }
    int numActive() const { return (int)active_.size(); }
    std::vector<int> takeDone();
    const Result &result(int id) const;
    void release(int id);
    size_t memoryInUse() const { return memInUse_; }
    size_t peakMemory() const { return memPeak_; }
    double now() const;
    static size_t estimateDenseBytes(int N, int m, int fieldFormat);
    static size_t estimateBipartiteBytes(int N0, int N1, int m, int fieldFormat);

private:
    struct Job;
    void admit();
    void finish(Job &j);
    void fail(Job &j, const char *what);
    void stepDenseGroup(const std::vector<DenseGraphAnnealer<float> *> &dense, const std::vector<Job *> &jobs,
                        const std::vector<int> &order, GroupedDenseStepper &stepper);
    bool denseGroupFits(int maxNp, int maxLd8, int maxM) const;
    void prepLoop();
    void enqueuePrep(Job *j);

    Device *dev_;
    std::vector<std::unique_ptr<Device>> pool_;
    size_t budget_ = (size_t)64 << 30;
    int maxActive_ = 256;
    int denseFormat_ = fieldBF16x2, bipartiteFormat_ = fieldBF16x2;
    double blockFraction_ = 1.0;
    int denseSpinLimit_ = 0; /* 0: the annealer's default */
    int nextId_ = 0, nextStream_ = 0;
    size_t memInUse_ = 0, memPeak_ = 0;
    std::vector<std::unique_ptr<Job>> pending_, active_;
    std::map<int, Result> done_;
    std::vector<int> newlyDone_;
    std::map<int, GroupedDenseStepper> steppers_; /* per size class */
    double t0_;
    /* host preparation (the jobs' Hamiltonians) on threads of their own,
     * started in submission order, so that a tick does not wait on it and
     * small jobs are not held behind a large one's preparation */
    std::mutex prepMu_;
    std::condition_variable prepCv_;
    std::deque<Job *> prepQueue_;
    bool prepStop_ = false;
    std::vector<std::thread> prepThreads_;
};

}  // namespace sqaod_rocm
