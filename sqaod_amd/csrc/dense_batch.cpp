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
#include <cstring>
#include <numeric>

#include "solver.h"

namespace sqaod_rocm {

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
    order_.resize(inst_.size());
    std::iota(order_.begin(), order_.end(), 0);
    if (largestFirst_)
        std::stable_sort(order_.begin(), order_.end(), [&](int x, int y) {
            const auto &a = *inst_[x], &b = *inst_[y];
            return a.getProblemSize() != b.getProblemSize() ? a.getProblemSize() > b.getProblemSize()
                                                            : a.numTrotters() > b.numTrotters();
        });
    maxNp_ = maxLd8_ = maxM_ = 0;
    for (auto *a : inst_) {
        maxNp_ = std::max(maxNp_, a->paddedN());
        maxLd8_ = std::max(maxLd8_, a->fp8RowStrideBytes());
        maxM_ = std::max(maxM_, a->numTrotters());
    }
    itemsValid_ = true;
}

void DenseBatch::anneal(const double *G, const double *beta, int nSteps, bool graph) {
    run(G, beta, nSteps, graph, wavesPerBlock_ > 0 ? wavesPerBlock_ : 1);
}

void DenseBatch::run(const double *G, const double *beta, int nSteps, bool graph, int wpb) {
    SQ_THROW_IF(inst_.empty(), "the batch holds no instances");
    if (!itemsValid_)
        buildItems();
    GroupedDenseStepper::Options o;
    o.fieldFormat = fieldFormat_;
    o.wavesPerBlock = wpb;
    o.graph = graph;
    stepper_.run(*dev_, inst_, order_, nSteps, [&](int k, int) { return std::make_pair(G[k], beta[k]); }, o);
    lastMs_ = stepper_.lastMs();
}

/* ---------------- GroupedDenseStepper ---------------- */

void GroupedDenseStepper::run(Device &dev, const std::vector<DenseGraphAnnealer<float> *> &inst,
                              const std::vector<int> &order, int nSteps, const Schedule &sched, const Options &o) {
    lastAborted_.clear();
    SQ_THROW_IF(inst.empty() || order.empty(), "grouped step: no instances");
    SQ_THROW_IF(nSteps <= 0, "no steps");
    dev.makeCurrent();
    (void)dev.soleSweeper(); /* its persistent sweeps spin: registered, so other Devices' sweeps launch cooperatively */
    const int nI = (int)inst.size();
    int maxNp = 0, maxLd8 = 0, maxM = 0;
    std::vector<uint2> items;
    for (int i : order) {
        SQ_THROW_IF(i < 0 || i >= nI, "grouped step: bad instance order");
        auto &a = *inst[i];
        SQ_THROW_IF(a.fieldFormat() != o.fieldFormat, "instance %d has another field format than the group", i);
        for (int y = 0; y < a.numTrotters(); ++y)
            items.push_back(uint2{(unsigned)i, (unsigned)y});
        maxNp = std::max(maxNp, a.paddedN());
        maxLd8 = std::max(maxLd8, a.fp8RowStrideBytes());
        maxM = std::max(maxM, a.numTrotters());
    }
    const int nItems = (int)items.size();
    std::vector<char> member(nI, 0);
    for (int i : order)
        member[i] = 1;
    for (int i : order) {
        inst[i]->throwErrorIfQNotSet();
        inst[i]->checkStepAbort();
        inst[i]->reserveTags(nSteps);
    }
    const int wpb = o.wavesPerBlock;
    const size_t regionBytes = denseSweepBatchRegionBytes(maxNp, maxLd8, sizeof(float), false, o.fieldFormat);
    SQ_THROW_IF(regionBytes * wpb + 256 > dev.maxLdsPerBlock(), "grouped step: N=%d too large for the LDS-resident sweep",
                maxNp);
    int capacity = denseSweepBatchBlocks<float>(regionBytes, wpb, false, o.fieldFormat, dev.numCUs());
    if (o.blockFraction > 0. && o.blockFraction < 1.)
        capacity = std::max(1, (int)(capacity * o.blockFraction));
    SQ_THROW_IF((long long)capacity * wpb < maxM, "grouped step: %d resident waves cannot hold an instance of %d slices",
                capacity * wpb, maxM);
    const int nBlocks = std::min(capacity, (nItems + wpb - 1) / wpb);

    /* plan every step on the host, in the order annealOneStep takes them */
    struct StepLayout {
        size_t gj, gs, sj, ss, sa;
        int nGemm, gemmBlocks, streamBlocks;
    };
    std::vector<StepLayout> lay(nSteps);
    std::vector<FieldGemmJob> gj;
    std::vector<int> gs, ss;
    std::vector<AttemptStreamJob> sj;
    std::vector<SweepArgs<float>> sa;
    for (int k = 0; k < nSteps; ++k) {
        StepLayout &L = lay[k];
        L.gj = gj.size();
        L.gs = gs.size();
        L.sj = sj.size();
        L.ss = ss.size();
        L.sa = sa.size();
        int gBlocks = 0, sBlocks = 0;
        for (int i = 0; i < nI; ++i) {
            auto &a = *inst[i];
            if (!member[i]) {
                sa.push_back(SweepArgs<float>{}); /* never referenced by an item */
                continue;
            }
            if (a.refreshDue()) {
                FieldGemmJob j = a.takeFieldGemmJob();
                gs.push_back(gBlocks);
                gBlocks += fieldGemmJobBlocks(j, o.fieldFormat);
                gj.push_back(j);
            }
            AttemptStreamJob s = a.takeAttemptStreamJob();
            ss.push_back(sBlocks);
            sBlocks += (int)(((long long)s.N * s.m + kAttemptBlock - 1) / kAttemptBlock);
            sj.push_back(s);
            const auto gb = sched(k, i);
            sa.push_back(a.makeSweepArgs((float)gb.first, (float)gb.second));
            a.stepDone();
        }
        L.nGemm = (int)(gj.size() - L.gj);
        L.gemmBlocks = gBlocks;
        L.streamBlocks = sBlocks;
    }
    const int nStream = (int)(sj.size() / nSteps);

    /* one pinned buffer, one upload */
    auto alignUp = [](size_t v) { return (v + 255) / 256 * 256; };
    size_t off = 0;
    const size_t oIT = off;
    off = alignUp(off + items.size() * sizeof(uint2));
    const size_t oGJ = off;
    off = alignUp(off + gj.size() * sizeof(FieldGemmJob));
    const size_t oGS = off;
    off = alignUp(off + gs.size() * sizeof(int));
    const size_t oSJ = off;
    off = alignUp(off + sj.size() * sizeof(AttemptStreamJob));
    const size_t oSS = off;
    off = alignUp(off + ss.size() * sizeof(int));
    const size_t oSA = off;
    off = alignUp(off + sa.size() * sizeof(SweepArgs<float>));
    h_plan_.allocate(off);
    if (d_plan_.n < off)
        d_plan_.allocate(off);
    d_queue_.allocate(1);
    unsigned char *h = h_plan_.ptr;
    memcpy(h + oIT, items.data(), items.size() * sizeof(uint2));
    if (!gj.empty())
        memcpy(h + oGJ, gj.data(), gj.size() * sizeof(FieldGemmJob));
    if (!gs.empty())
        memcpy(h + oGS, gs.data(), gs.size() * sizeof(int));
    memcpy(h + oSJ, sj.data(), sj.size() * sizeof(AttemptStreamJob));
    memcpy(h + oSS, ss.data(), ss.size() * sizeof(int));
    memcpy(h + oSA, sa.data(), sa.size() * sizeof(SweepArgs<float>));
    hipStream_t st = dev.stream();
    HIP_CHECK(hipMemcpyAsync(d_plan_.ptr, h, off, hipMemcpyHostToDevice, st));
    unsigned char *d = d_plan_.ptr;

    auto launchAll = [&]() {
        for (int k = 0; k < nSteps; ++k) {
            const StepLayout &L = lay[k];
            if (L.nGemm > 0)
                fieldRefreshGrouped(o.fieldFormat, reinterpret_cast<const FieldGemmJob *>(d + oGJ) + L.gj,
                                    reinterpret_cast<const int *>(d + oGS) + L.gs, L.nGemm, L.gemmBlocks, st);
            philoxAttemptsGrouped<float>(reinterpret_cast<const AttemptStreamJob *>(d + oSJ) + L.sj,
                                         reinterpret_cast<const int *>(d + oSS) + L.ss, nStream, L.streamBlocks, st);
            denseSweepBatch<float>(reinterpret_cast<const SweepArgs<float> *>(d + oSA) + L.sa,
                                   reinterpret_cast<const uint2 *>(d + oIT), nItems, d_queue_.ptr, false,
                                   o.fieldFormat, regionBytes, wpb, nBlocks, st);
        }
    };

    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    auto cleanup = [&]() {
        if (ge)
            (void)hipGraphExecDestroy(ge);
        if (g)
            (void)hipGraphDestroy(g);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    };
    try {
        if (o.graph) {
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
        if (o.wait) {
            HIP_CHECK(hipStreamSynchronize(st));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            lastMs_ = ms;
        }
    } catch (...) {
        cleanup();
        throw;
    }
    if (!o.wait)
        HIP_CHECK(hipStreamSynchronize(st)); /* the plan buffer and graph are released below */
    cleanup();
    std::string first;
    for (int i : order) {
        try {
            inst[i]->checkStepAbort();
        } catch (const std::exception &e) {
            lastAborted_.push_back(i);
            if (first.empty())
                first = e.what();
        }
    }
    SQ_THROW_IF(!lastAborted_.empty(), "%s (%zu instance(s))", first.c_str(), lastAborted_.size());
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
