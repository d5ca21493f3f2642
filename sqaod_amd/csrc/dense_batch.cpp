/*
 * Batch engine: many dense annealers stepped together.  The reference has
 * no batched solver; each instance here is a DenseGraphAnnealer<float>
 * (setQUBO, prepare, randomizeSpin, get_E/get_x as CPUDenseGraphAnnealer /
 * CUDADenseGraphAnnealer), and a step of the batch issues, for all
 * instances at once, the launches annealOneStep issues for one:
 *
 *   grouped field GEMM (instances whose refresh is due)  -> fieldRefreshGrouped
 *   grouped Philox attempt streams                        -> philoxAttemptsGrouped
 *   one persistent sweep over every (instance, slice)     -> denseSweepBatch
 *
 * with the per-instance arguments built by the annealer's own helpers
 * (makeSweepArgs, takeFieldGemmJob, takeAttemptStreamJob), so every instance
 * follows exactly the trajectory it would follow alone.  A schedule of K
 * steps is planned on the host, uploaded in one copy, and either launched
 * directly or captured into one hipGraph and launched once.
 */
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include "solver.h"

namespace sqaod_rocm {

namespace {

size_t alignUp(size_t v, size_t a) { return (v + a - 1) / a * a; }

/* offsets of one step's arrays inside the uploaded plan */
struct StepLayout {
    size_t gemmJobs, gemmStart, streamJobs, streamStart, args;
    int nGemm, gemmBlocks, streamBlocks;
};

}  // namespace

DenseBatch::DenseBatch(Device &dev) : dev_(&dev) {}

DenseBatch::~DenseBatch() {
    for (auto *a : inst_)
        delete a;
}

DenseGraphAnnealer<float> &DenseBatch::instance(int i) {
    SQ_THROW_IF(i < 0 || i >= (int)inst_.size(), "instance %d out of range (%zu instances)", i, inst_.size());
    return *inst_[i];
}

void DenseBatch::setPreference(const std::string &name, const std::string &value) {
    if (name == "waves_per_block") {
        const int w = atoi(value.c_str());
        SQ_THROW_IF(w != 0 && w != 1 && w != 2 && w != 4, "waves_per_block must be 0 (tuned), 1, 2 or 4");
        wavesPerBlock_ = w;
        return;
    }
    if (name == "order") {
        SQ_THROW_IF(value != "largest_first" && value != "given", "order must be largest_first or given");
        largestFirst_ = value == "largest_first";
        itemsValid_ = false;
        return;
    }
    SQ_THROW_IF(name != "field_precision" && name != "refresh_interval" && name != "spin_limit" &&
                    name != "algorithm",
                "unknown batch preference '%s'", name.c_str());
    for (auto *a : inst_)
        a->setPreference(name, value);
    if (name == "field_precision") {
        DenseGraphAnnealer<float> probe;
        probe.setPreference(name, value); /* validates and maps the name */
        fieldFormat_ = probe.fieldFormat();
    }
    for (auto &p : prefs_)
        if (p.first == name) {
            p.second = value;
            return;
        }
    prefs_.emplace_back(name, value);
}

PrefList DenseBatch::getPreferences() const {
    PrefList p = prefs_;
    p.emplace_back("waves_per_block", std::to_string(wavesPerBlock_));
    p.emplace_back("order", largestFirst_ ? "largest_first" : "given");
    p.emplace_back("instances", std::to_string(inst_.size()));
    return p;
}

int DenseBatch::addQUBO(const float *W, int N, int m, unsigned long long seed, OptimizeMethod om) {
    SQ_THROW_IF(m <= 0, "# trotters must be a positive integer.");
    auto *a = new DenseGraphAnnealer<float>();
    try {
        a->assignDevice(*dev_);
        a->seed(seed);
        a->setQUBO(W, N, om);
        a->setPreference("n_trotters", std::to_string(m));
        for (auto &p : prefs_)
            a->setPreference(p.first, p.second);
        std::string why;
        SQ_THROW_IF(!a->batchable(&why), "instance cannot join a batch: %s", why.c_str());
    } catch (...) {
        delete a;
        throw;
    }
    inst_.push_back(a);
    itemsValid_ = false;
    return (int)inst_.size() - 1;
}

void DenseBatch::prepare() {
    SQ_THROW_IF(inst_.empty(), "the batch holds no instances");
    for (auto *a : inst_) {
        a->prepare();
        a->randomizeSpin();
    }
    itemsValid_ = false;
    buildItems();
}

long long DenseBatch::attemptsPerStep() const {
    long long s = 0;
    for (auto *a : inst_)
        s += (long long)a->getProblemSize() * a->numTrotters();
    return s;
}

void DenseBatch::buildItems() {
    std::vector<int> order(inst_.size());
    std::iota(order.begin(), order.end(), 0);
    if (largestFirst_)
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            const auto &a = *inst_[x], &b = *inst_[y];
            return a.getProblemSize() != b.getProblemSize() ? a.getProblemSize() > b.getProblemSize()
                                                            : a.numTrotters() > b.numTrotters();
        });
    std::vector<uint2> items;
    maxNp_ = maxLd8_ = maxM_ = 0;
    for (int i : order) {
        auto &a = *inst_[i];
        SQ_THROW_IF(a.fieldFormat() != fieldFormat_, "instance %d has another field format than the batch", i);
        for (int y = 0; y < a.numTrotters(); ++y)
            items.push_back(uint2{(unsigned)i, (unsigned)y});
        maxNp_ = std::max(maxNp_, a.paddedN());
        maxLd8_ = std::max(maxLd8_, a.fp8RowStrideBytes());
        maxM_ = std::max(maxM_, a.numTrotters());
    }
    nItems_ = (int)items.size();
    d_items_.allocate(items.size());
    HIP_CHECK(hipMemcpy(d_items_.ptr, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice));
    d_queue_.allocate(1);
    itemsValid_ = true;
}

void DenseBatch::anneal(const double *G, const double *beta, int nSteps, bool graph) {
    run(G, beta, nSteps, graph, wavesPerBlock_ > 0 ? wavesPerBlock_ : 1);
}

void DenseBatch::run(const double *G, const double *beta, int nSteps, bool graph, int wpb) {
    SQ_THROW_IF(inst_.empty(), "the batch holds no instances");
    SQ_THROW_IF(nSteps <= 0, "no steps");
    dev_->makeCurrent();
    if (!itemsValid_)
        buildItems();
    const int nI = (int)inst_.size();
    for (auto *a : inst_) {
        a->throwErrorIfQNotSet();
        a->checkStepAbort();
        a->reserveTags(nSteps);
    }
    const size_t regionBytes = denseSweepBatchRegionBytes(maxNp_, maxLd8_, sizeof(float), false, fieldFormat_);
    const size_t lds = regionBytes * wpb + 256;
    SQ_THROW_IF(lds > dev_->maxLdsPerBlock(), "batch: N=%d too large for the LDS-resident sweep", maxNp_);
    const int capacity = denseSweepBatchBlocks<float>(regionBytes, wpb, false, fieldFormat_, dev_->numCUs());
    SQ_THROW_IF((long long)capacity * wpb < maxM_,
                "batch: %d resident waves cannot hold an instance of %d slices", capacity * wpb, maxM_);
    const int nBlocks = std::min(capacity, (nItems_ + wpb - 1) / wpb);

    /* plan every step on the host, in the order annealOneStep takes them */
    std::vector<StepLayout> lay(nSteps);
    std::vector<FieldGemmJob> gj;
    std::vector<int> gs;
    std::vector<AttemptStreamJob> sj;
    std::vector<int> ss;
    std::vector<SweepArgs<float>> sa;
    std::vector<size_t> gjOff(nSteps), sjOff(nSteps), saOff(nSteps), gsOff(nSteps), ssOff(nSteps);
    for (int k = 0; k < nSteps; ++k) {
        StepLayout &L = lay[k];
        gjOff[k] = gj.size();
        gsOff[k] = gs.size();
        sjOff[k] = sj.size();
        ssOff[k] = ss.size();
        saOff[k] = sa.size();
        int gBlocks = 0, sBlocks = 0;
        for (int i = 0; i < nI; ++i) {
            auto &a = *inst_[i];
            if (a.refreshDue()) {
                FieldGemmJob j = a.takeFieldGemmJob();
                gs.push_back(gBlocks);
                gBlocks += j.nCols / 16;
                gj.push_back(j);
            }
            AttemptStreamJob s = a.takeAttemptStreamJob();
            ss.push_back(sBlocks);
            sBlocks += (int)(((long long)s.N * s.m + kAttemptBlock - 1) / kAttemptBlock);
            sj.push_back(s);
            sa.push_back(a.makeSweepArgs((float)G[k], (float)beta[k]));
            a.stepDone();
        }
        L.nGemm = (int)(gj.size() - gjOff[k]);
        L.gemmBlocks = gBlocks;
        L.streamBlocks = sBlocks;
    }
    /* one pinned buffer, one upload */
    size_t off = 0;
    const size_t oGJ = off;
    off = alignUp(off + gj.size() * sizeof(FieldGemmJob), 256);
    const size_t oGS = off;
    off = alignUp(off + gs.size() * sizeof(int), 256);
    const size_t oSJ = off;
    off = alignUp(off + sj.size() * sizeof(AttemptStreamJob), 256);
    const size_t oSS = off;
    off = alignUp(off + ss.size() * sizeof(int), 256);
    const size_t oSA = off;
    off = alignUp(off + sa.size() * sizeof(SweepArgs<float>), 256);
    h_plan_.allocate(off);
    if (d_plan_.n < off)
        d_plan_.allocate(off);
    unsigned char *h = h_plan_.ptr;
    if (!gj.empty())
        memcpy(h + oGJ, gj.data(), gj.size() * sizeof(FieldGemmJob));
    if (!gs.empty())
        memcpy(h + oGS, gs.data(), gs.size() * sizeof(int));
    memcpy(h + oSJ, sj.data(), sj.size() * sizeof(AttemptStreamJob));
    memcpy(h + oSS, ss.data(), ss.size() * sizeof(int));
    memcpy(h + oSA, sa.data(), sa.size() * sizeof(SweepArgs<float>));
    hipStream_t st = dev_->stream();
    HIP_CHECK(hipMemcpyAsync(d_plan_.ptr, h, off, hipMemcpyHostToDevice, st));
    unsigned char *d = d_plan_.ptr;

    auto launchAll = [&]() {
        for (int k = 0; k < nSteps; ++k) {
            const StepLayout &L = lay[k];
            if (L.nGemm > 0)
                fieldRefreshGrouped(fieldFormat_, reinterpret_cast<const FieldGemmJob *>(d + oGJ) + gjOff[k],
                                    reinterpret_cast<const int *>(d + oGS) + gsOff[k], L.nGemm, L.gemmBlocks, st);
            philoxAttemptsGrouped<float>(reinterpret_cast<const AttemptStreamJob *>(d + oSJ) + sjOff[k],
                                         reinterpret_cast<const int *>(d + oSS) + ssOff[k], nI, L.streamBlocks, st);
            denseSweepBatch<float>(reinterpret_cast<const SweepArgs<float> *>(d + oSA) + saOff[k], d_items_.ptr,
                                   nItems_, d_queue_.ptr, false, fieldFormat_, regionBytes, wpb, nBlocks, st);
        }
    };

    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    try {
        if (graph) {
            HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
            try {
                launchAll();
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                throw;
            }
            HIP_CHECK(hipStreamEndCapture(st, &g));
            HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            HIP_CHECK(hipEventRecord(e0, st));
            HIP_CHECK(hipGraphLaunch(ge, st));
            HIP_CHECK(hipEventRecord(e1, st));
        } else {
            HIP_CHECK(hipEventRecord(e0, st));
            launchAll();
            HIP_CHECK(hipEventRecord(e1, st));
        }
        HIP_CHECK(hipStreamSynchronize(st));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        lastMs_ = ms;
    } catch (...) {
        if (ge)
            (void)hipGraphExecDestroy(ge);
        if (g)
            (void)hipGraphDestroy(g);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    if (ge)
        HIP_CHECK(hipGraphExecDestroy(ge));
    if (g)
        HIP_CHECK(hipGraphDestroy(g));
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    for (auto *a : inst_)
        a->checkStepAbort();
}

void DenseBatch::autotune(double G, double beta, int trialSteps) {
    SQ_THROW_IF(inst_.empty(), "the batch holds no instances");
    SQ_THROW_IF(trialSteps <= 0, "autotune needs at least one trial step");
    dev_->makeCurrent();
    /* snapshot spins, fields and stream positions */
    std::vector<size_t> qOff, fOff;
    size_t qTot = 0, fTot = 0;
    for (auto *a : inst_) {
        qOff.push_back(qTot);
        fOff.push_back(fTot);
        qTot += (size_t)a->numTrotters() * (a->paddedN() / 32);
        fTot += (size_t)a->numTrotters() * a->paddedN();
    }
    DevBuf<uint32_t> snapQ;
    DevBuf<float> snapF;
    snapQ.allocate(qTot);
    snapF.allocate(fTot);
    std::vector<DenseGraphAnnealer<float>::StepState> ss;
    hipStream_t st = dev_->stream();
    for (size_t i = 0; i < inst_.size(); ++i) {
        auto *a = inst_[i];
        const size_t nq = (size_t)a->numTrotters() * (a->paddedN() / 32), nf = (size_t)a->numTrotters() * a->paddedN();
        HIP_CHECK(hipMemcpyAsync(snapQ.ptr + qOff[i], a->spinWords(), nq * 4, hipMemcpyDeviceToDevice, st));
        HIP_CHECK(hipMemcpyAsync(snapF.ptr + fOff[i], a->fields(), nf * 4, hipMemcpyDeviceToDevice, st));
        ss.push_back(a->stepState());
    }
    auto restore = [&]() {
        for (size_t i = 0; i < inst_.size(); ++i) {
            auto *a = inst_[i];
            const size_t nq = (size_t)a->numTrotters() * (a->paddedN() / 32),
                         nf = (size_t)a->numTrotters() * a->paddedN();
            HIP_CHECK(hipMemcpyAsync(a->spinWords(), snapQ.ptr + qOff[i], nq * 4, hipMemcpyDeviceToDevice, st));
            HIP_CHECK(hipMemcpyAsync(a->fields(), snapF.ptr + fOff[i], nf * 4, hipMemcpyDeviceToDevice, st));
            a->setStepState(ss[i]);
        }
        HIP_CHECK(hipStreamSynchronize(st));
    };
    std::vector<double> Gs(trialSteps, G), Bs(trialSteps, beta);
    double best = 1e300;
    int bestW = 1;
    bool bestOrder = true;
    const bool savedOrder = largestFirst_;
    for (int order = 0; order < 2; ++order) {
        largestFirst_ = order == 0;
        itemsValid_ = false;
        buildItems();
        for (int w : {1, 2, 4}) {
            const size_t regionBytes = denseSweepBatchRegionBytes(maxNp_, maxLd8_, sizeof(float), false, fieldFormat_);
            if (regionBytes * w + 256 > dev_->maxLdsPerBlock())
                continue;
            const int cap = denseSweepBatchBlocks<float>(regionBytes, w, false, fieldFormat_, dev_->numCUs());
            if ((long long)cap * w < maxM_)
                continue;
            run(Gs.data(), Bs.data(), trialSteps, false, w);
            restore();
            logf("batch autotune: waves_per_block=%d order=%s: %.3f ms", w, largestFirst_ ? "largest_first" : "given",
                 lastMs_);
            if (lastMs_ < best) {
                best = lastMs_;
                bestW = w;
                bestOrder = largestFirst_;
            }
        }
    }
    (void)savedOrder;
    wavesPerBlock_ = bestW;
    largestFirst_ = bestOrder;
    itemsValid_ = false;
    buildItems();
}

}  // namespace sqaod_rocm
