/*
 * Serving: continuous batching of anneal jobs on one device.  The reference
 * has no server; the jobs are the reference's solvers (a dense job is a
 * DenseGraphAnnealer<float>, a bipartite job a BipartiteGraphAnnealer<float>,
 * Solver.h:187-252), each driven through its own schedule of
 * annealOneStep(G, beta) calls and finished with makeSolution / get_E /
 * get_x.  What the server adds is when and together with what a step runs:
 *
 *  - admission in arrival order while the active-job limit and the device
 *    memory budget (an estimate of each job's device footprint) allow;
 *  - one tick = one step of every active job: bipartite steps are enqueued
 *    on a pool of streams, then the dense jobs run through the grouped
 *    stepper (grouped field GEMM + grouped Philox streams + one persistent
 *    sweep) on the server's stream, one launch per size class so small jobs
 *    are not sized for the largest; the streams are joined at the end;
 *  - a job whose schedule is done is finished, its device memory is freed
 *    and the next pending jobs can be admitted at the following tick.
 *
 * A job's host preparation (its Hamiltonian, O(N^2) on the host) runs on one
 * of a few threads of the server, started in submission order; admission
 * (still in arrival order) waits for it only when it has not finished, so a
 * large job's preparation neither stalls the ticks of the jobs already
 * running nor delays the preparation of the jobs behind it.
 *
 * Only one persistent sweep runs at a time (size classes are serialised on
 * one stream): two persistent launches that were each partly resident could
 * wait on each other's resources.  Bipartite kernels never wait, so they can
 * share the device with it.
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>

#include "solver.h"

namespace sqaod_rocm {

struct Server::Job {
    int id = 0, kind = kDense, N0 = 0, N1 = 0, m = 0;
    unsigned long long seed = 0;
    std::vector<float> W, b0, b1; /* host copies until the preparation */
    /* the preparation thread's result: 0 queued, 1 ready, 2 failed */
    std::atomic<int> prep{0};
    std::string prepError;
    std::vector<float> h, J; /* dense Hamiltonian */
    float c = 0.f;
    std::vector<double> h0d, h1d, Jd; /* bipartite Hamiltonian */
    double cd = 0.;
    int spinLimit = 0; /* dense: the server's dense_spin_limit when submitted */
    std::vector<double> G, beta;
    int step = 0;
    size_t bytes = 0;
    bool admitted = false, failed = false;
    double submitS = 0., admitS = 0.;
    std::unique_ptr<DenseGraphAnnealer<float>> dense;
    std::unique_ptr<BipartiteGraphAnnealer<float>> bip;
};

namespace {

const char *formatName(int f) { return f == fieldBF16x2 ? "bf16x2" : f == fieldFP8 ? "fp8" : "fp32"; }

int formatOf(const std::string &v) {
    if (v == "fp32")
        return fieldF32;
    if (v == "bf16x2")
        return fieldBF16x2;
    if (v == "fp8")
        return fieldFP8;
    SQ_THROW("unknown field precision '%s' (fp32 | bf16x2 | fp8)", v.c_str());
}

size_t roundUp256(int n) { return (size_t)((n + 255) / 256 * 256); }

constexpr int kPrepThreads = 4; /* threads forming the jobs' Hamiltonians */

/* dense jobs share a persistent sweep with jobs of the same padded-size class */
int sizeClass(int Np) {
    int c = 1024;
    while (c < Np)
        c <<= 1;
    return c;
}

}  // namespace

Server::Server(Device &dev) : dev_(&dev) {
    t0_ = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    for (int i = 0; i < 4; ++i)
        pool_.emplace_back(new Device(dev.devno()));
    for (int i = 0; i < kPrepThreads; ++i)
        prepThreads_.emplace_back([this] { prepLoop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(prepMu_);
        prepStop_ = true;
    }
    prepCv_.notify_all();
    for (auto &t : prepThreads_)
        t.join();
    active_.clear();
    pending_.clear();
}

/* A preparation thread: the queued jobs' Hamiltonians (the host half of
 * setQUBO), taken in submission order.  A job stays in pending_ (owned by the
 * server thread) until admission, which takes it only after prep != 0, so
 * the job outlives its preparation. */
void Server::prepLoop() {
    for (;;) {
        Job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(prepMu_);
            prepCv_.wait(lk, [this] { return prepStop_ || !prepQueue_.empty(); });
            if (prepStop_)
                return;
            j = prepQueue_.front();
            prepQueue_.pop_front();
        }
        try {
            if (j->kind == kDense) {
                /* W was checked for symmetry at submission */
                DenseGraphAnnealer<float>::hamiltonianOf(j->W.data(), j->N0, optMinimize, false, j->h, j->J, j->c);
            } else {
                BipartiteGraphAnnealer<float>::hamiltonianOf(j->b0.data(), j->b1.data(), j->W.data(), j->N0, j->N1,
                                                             optMinimize, j->h0d, j->h1d, j->Jd, j->cd);
            }
            std::vector<float>().swap(j->W);
            std::vector<float>().swap(j->b0);
            std::vector<float>().swap(j->b1);
            j->prep.store(1, std::memory_order_release);
        } catch (const std::exception &e) {
            j->prepError = e.what();
            j->prep.store(2, std::memory_order_release);
        }
    }
}

void Server::enqueuePrep(Job *j) {
    {
        std::lock_guard<std::mutex> lk(prepMu_);
        prepQueue_.push_back(j);
    }
    prepCv_.notify_one();
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t0_;
}

void Server::setPreference(const std::string &name, const std::string &value) {
    if (name == "memory_budget_mb") {
        const long long mb = atoll(value.c_str());
        SQ_THROW_IF(mb <= 0, "memory_budget_mb must be positive");
        budget_ = (size_t)mb << 20;
    } else if (name == "max_active") {
        maxActive_ = atoi(value.c_str());
        SQ_THROW_IF(maxActive_ <= 0, "max_active must be positive");
    } else if (name == "dense_field_precision") {
        SQ_THROW_IF(!active_.empty() || !pending_.empty(), "set the field precision before submitting jobs");
        denseFormat_ = formatOf(value);
    } else if (name == "bipartite_field_precision") {
        SQ_THROW_IF(!active_.empty() || !pending_.empty(), "set the field precision before submitting jobs");
        bipartiteFormat_ = formatOf(value);
        SQ_THROW_IF(bipartiteFormat_ == fieldFP8, "bipartite jobs take fp32 or bf16x2");
    } else if (name == "bipartite_streams") {
        const int n = atoi(value.c_str());
        SQ_THROW_IF(n < 1 || n > 16, "bipartite_streams must be in 1..16");
        SQ_THROW_IF(!active_.empty(), "set the stream count while no job is active");
        pool_.clear();
        for (int i = 0; i < n; ++i)
            pool_.emplace_back(new Device(dev_->devno()));
        nextStream_ = 0;
    } else if (name == "block_fraction") {
        blockFraction_ = atof(value.c_str());
        SQ_THROW_IF(!(blockFraction_ > 0. && blockFraction_ <= 1.), "block_fraction must be in (0, 1]");
    } else if (name == "dense_spin_limit") {
        /* the sweep's bounded wait of the dense jobs submitted from now on
         * (0: the annealer's default); a tiny limit makes their steps fail */
        denseSpinLimit_ = atoi(value.c_str());
        SQ_THROW_IF(denseSpinLimit_ < 0, "dense_spin_limit must be >= 0");
    } else {
        SQ_THROW("unknown server preference '%s'", name.c_str());
    }
}

size_t Server::estimateDenseBytes(int N, int m, int fieldFormat) {
    const size_t Np = roundUp256(N), n = (size_t)N;
    size_t b = n * Np * 4 + Np * 4;                                 /* J, h */
    if (fieldFormat == fieldBF16x2)
        b += n * Np * 4 + Np * Np * 4;                              /* Jsplit, hi/lo fragments */
    else if (fieldFormat == fieldFP8)
        b += n * Np * 4 + n * ((Np + 1023) / 1024 * 1024) + Np * Np + Np * 8; /* J', rows, fragments, scales */
    else
        b += Np * Np * 4;                                           /* fp32 fragments */
    b += (size_t)m * Np * 4 + 3 * (size_t)m * n * 4 + (size_t)m * Np / 8; /* F, X, U, D, Q */
    b += (size_t)m * (Np / 256 + 2) * 8 + (size_t)m * n;            /* energies, q for makeSolution */
    b += 6 * Np * Np + (size_t)m * (2048 + Np / 16) * 8;            /* energy planes (at most 6), partials */
    return b + b / 16 + ((size_t)1 << 20);
}

size_t Server::estimateBipartiteBytes(int N0, int N1, int m, int fieldFormat) {
    const size_t N0p = roundUp256(N0), N1p = roundUp256(N1);
    size_t b = (size_t)N1 * N0p * 4 + (size_t)N0 * N1p * 4 + (N0p + N1p) * 4; /* J, J^T, h0, h1 */
    b += 2 * N0p * N1p * 4;                               /* fragments of J and J^T (either form) */
    (void)fieldFormat;
    b += (size_t)m * std::max(N0p, N1p) * 4;              /* dEmat */
    b += (size_t)m * (N0p + N1p) / 8;                     /* spins */
    b += (size_t)m * ((N0p + N1p) / 256 + 2) * 8 + (size_t)m * (N0 + N1);
    b += 6 * N0p * N1p + (size_t)m * (2048 + N1p / 16) * 8;         /* energy planes (at most 6), partials */
    return b + b / 16 + ((size_t)1 << 20);
}

int Server::submitDense(const float *W, int N, int m, unsigned long long seed, const double *G, const double *beta,
                        int nSteps) {
    SQ_THROW_IF(!W || N <= 0 || m <= 0 || nSteps <= 0 || !G || !beta, "bad dense job");
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < i; ++j)
            SQ_THROW_IF(W[(size_t)i * N + j] != W[(size_t)j * N + i], "W is not symmetric.");
    auto j = std::make_unique<Job>();
    j->id = nextId_++;
    j->kind = kDense;
    j->N0 = N;
    j->m = m;
    j->seed = seed;
    j->W.assign(W, W + (size_t)N * N);
    j->G.assign(G, G + nSteps);
    j->beta.assign(beta, beta + nSteps);
    j->bytes = estimateDenseBytes(N, m, denseFormat_);
    j->spinLimit = denseSpinLimit_;
    SQ_THROW_IF(j->bytes > budget_, "job needs ~%zu MiB, over the %zu MiB budget", j->bytes >> 20, budget_ >> 20);
    /* the grouped sweep holds a job's fields in one wave's LDS region and
     * needs its m slices resident at once: a job past either limit could
     * never be stepped */
    const int Np = (int)roundUp256(N);
    SQ_THROW_IF(!denseGroupFits(Np, fp8RowStride(Np), m),
                "dense job N=%d, m=%d does not fit the grouped sweep (LDS region per slice or resident waves)", N,
                m);
    j->submitS = now();
    enqueuePrep(j.get());
    pending_.push_back(std::move(j));
    return pending_.back()->id;
}

int Server::submitBipartite(const float *b0, const float *b1, const float *W, int N0, int N1, int m,
                            unsigned long long seed, const double *G, const double *beta, int nSteps) {
    SQ_THROW_IF(!b0 || !b1 || !W || N0 <= 0 || N1 <= 0 || m <= 0 || nSteps <= 0 || !G || !beta, "bad bipartite job");
    auto j = std::make_unique<Job>();
    j->id = nextId_++;
    j->kind = kBipartite;
    j->N0 = N0;
    j->N1 = N1;
    j->m = m;
    j->seed = seed;
    j->b0.assign(b0, b0 + N0);
    j->b1.assign(b1, b1 + N1);
    j->W.assign(W, W + (size_t)N1 * N0);
    j->G.assign(G, G + nSteps);
    j->beta.assign(beta, beta + nSteps);
    j->bytes = estimateBipartiteBytes(N0, N1, m, bipartiteFormat_);
    SQ_THROW_IF(j->bytes > budget_, "job needs ~%zu MiB, over the %zu MiB budget", j->bytes >> 20, budget_ >> 20);
    j->submitS = now();
    enqueuePrep(j.get());
    pending_.push_back(std::move(j));
    return pending_.back()->id;
}

bool Server::denseGroupFits(int maxNp, int maxLd8, int maxM) const {
    const size_t region = denseSweepBatchRegionBytes(maxNp, maxLd8, sizeof(float), false, denseFormat_);
    if (region + 256 > dev_->maxLdsPerBlock())
        return false;
    int capacity = denseSweepBatchBlocks<float>(region, 1, false, denseFormat_, dev_->numCUs());
    if (blockFraction_ < 1.)
        capacity = std::max(1, (int)(capacity * blockFraction_));
    return capacity >= maxM;
}

void Server::admit() {
    while (!pending_.empty()) {
        Job &j = *pending_.front();
        const int prep = j.prep.load(std::memory_order_acquire);
        if (prep == 0)
            break; /* arrival order: the head job's Hamiltonian is still being formed */
        if (prep == 1 && ((int)active_.size() >= maxActive_ || memInUse_ + j.bytes > budget_))
            break; /* arrival order: a job that does not fit holds the ones behind it */
        std::unique_ptr<Job> job = std::move(pending_.front());
        pending_.erase(pending_.begin());
        if (prep == 2) {
            fail(j, j.prepError.c_str());
            continue;
        }
        try {
            if (j.kind == kDense) {
                j.dense.reset(new DenseGraphAnnealer<float>());
                auto &a = *j.dense;
                a.assignDevice(*dev_);
                a.seed(j.seed);
                a.setProblem(j.h.data(), j.J.data(), j.c, j.N0, optMinimize);
                a.setPreference("n_trotters", std::to_string(j.m));
                a.setPreference("field_precision", formatName(denseFormat_));
                if (j.spinLimit > 0)
                    a.setPreference("spin_limit", std::to_string(j.spinLimit));
                a.prepare();
                a.randomizeSpin();
            } else {
                j.bip.reset(new BipartiteGraphAnnealer<float>());
                auto &a = *j.bip;
                a.assignDevice(*pool_[nextStream_]);
                nextStream_ = (nextStream_ + 1) % (int)pool_.size();
                a.seed(j.seed);
                a.setProblem(j.h0d, j.h1d, j.Jd, j.cd, j.N0, j.N1, optMinimize);
                a.setPreference("n_trotters", std::to_string(j.m));
                a.setPreference("field_precision", formatName(bipartiteFormat_));
                a.prepare();
                a.randomizeSpin();
            }
        } catch (const std::exception &e) {
            fail(j, e.what()); /* device state (if any) is freed with the job */
            continue;
        }
        std::vector<float>().swap(j.h);
        std::vector<float>().swap(j.J);
        std::vector<double>().swap(j.h0d);
        std::vector<double>().swap(j.h1d);
        std::vector<double>().swap(j.Jd);
        j.admitS = now();
        j.admitted = true;
        memInUse_ += j.bytes;
        memPeak_ = std::max(memPeak_, memInUse_);
        active_.push_back(std::move(job));
    }
}

void Server::stepDenseGroup(const std::vector<DenseGraphAnnealer<float> *> &dense, const std::vector<Job *> &jobs,
                            const std::vector<int> &order, GroupedDenseStepper &stepper) {
    GroupedDenseStepper::Options o;
    o.fieldFormat = denseFormat_;
    o.wavesPerBlock = 1;
    o.graph = false;
    o.wait = true;
    o.blockFraction = blockFraction_;
    try {
        stepper.run(*dev_, dense, order, 1,
                    [&](int, int i) {
                        const Job *jb = jobs[i];
                        return std::make_pair(jb->G[jb->step], jb->beta[jb->step]);
                    },
                    o);
    } catch (const std::exception &e) {
        /* jobs whose sweep aborted fail; if the step failed otherwise, none
         * of the group's jobs has a valid state and they all fail; the other
         * groups and the bipartite jobs go on */
        const std::vector<int> &ab = stepper.lastAborted();
        for (int i : (ab.empty() ? order : ab))
            if (!jobs[i]->failed)
                fail(*jobs[i], e.what());
    }
}

int Server::tick() {
    admit();
    if (active_.empty())
        return (int)pending_.size();
    /* bipartite steps first, on their own streams */
    for (auto &j : active_)
        if (j->kind == kBipartite) {
            try {
                j->bip->annealOneStep((float)j->G[j->step], (float)j->beta[j->step]);
            } catch (const std::exception &e) {
                fail(*j, e.what());
            }
        }
    /* dense steps: one persistent sweep per size class, on the server's stream */
    std::vector<DenseGraphAnnealer<float> *> dense;
    std::vector<Job *> denseJobs;
    std::map<int, std::vector<int>> classes;
    for (auto &j : active_)
        if (j->kind == kDense) {
            classes[sizeClass(j->dense->paddedN())].push_back((int)dense.size());
            dense.push_back(j->dense.get());
            denseJobs.push_back(j.get());
        }
    for (auto &c : classes) {
        std::vector<int> &order = c.second;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return dense[x]->getProblemSize() != dense[y]->getProblemSize()
                       ? dense[x]->getProblemSize() > dense[y]->getProblemSize()
                       : dense[x]->numTrotters() > dense[y]->numTrotters();
        });
        /* Every job fits the grouped sweep alone (checked at submission), but
         * the largest N and the largest m of a class may not fit one launch
         * together: such a class runs as consecutive groups, each as large
         * as fits (jobs in the class's order). */
        size_t g0 = 0;
        while (g0 < order.size()) {
            std::vector<int> group;
            int maxNp = 0, maxLd8 = 0, maxM = 0;
            for (size_t g = g0; g < order.size(); ++g) {
                const DenseGraphAnnealer<float> &a = *dense[order[g]];
                const int np = std::max(maxNp, a.paddedN()), l8 = std::max(maxLd8, a.fp8RowStrideBytes());
                const int mm = std::max(maxM, a.numTrotters());
                if (!group.empty() && !denseGroupFits(np, l8, mm))
                    break;
                group.push_back(order[g]);
                maxNp = np;
                maxLd8 = l8;
                maxM = mm;
            }
            g0 += group.size();
            stepDenseGroup(dense, denseJobs, group, steppers_[c.first]);
        }
    }
    for (auto &d : pool_)
        d->synchronize();
    /* advance; finish the jobs whose schedule is done.  active_ is rebuilt
     * from moved-out entries only after every job has been handled, and a job
     * whose solution throws becomes a failed result, so a failure leaves no
     * job half-moved. */
    std::vector<std::unique_ptr<Job>> all;
    all.swap(active_);
    for (auto &j : all) {
        if (j->failed)
            continue; /* its result is recorded */
        ++j->step;
        if (j->step >= (int)j->G.size()) {
            try {
                finish(*j);
            } catch (const std::exception &e) {
                fail(*j, e.what());
            }
        } else {
            active_.push_back(std::move(j));
        }
    }
    return (int)(active_.size() + pending_.size());
}

void Server::fail(Job &j, const char *what) {
    Result r;
    r.kind = j.kind;
    r.N0 = j.N0;
    r.N1 = j.N1;
    r.m = j.m;
    r.attempts = 0;
    r.submitS = j.submitS;
    r.admitS = j.admitS;
    r.finishS = now();
    r.error = (what && *what) ? what : "job failed";
    j.dense.reset();
    j.bip.reset();
    if (j.admitted) {
        memInUse_ -= j.bytes;
        j.admitted = false;
    }
    j.failed = true;
    done_[j.id] = std::move(r);
    newlyDone_.push_back(j.id);
}

void Server::finish(Job &j) {
    Result r;
    r.kind = j.kind;
    r.N0 = j.N0;
    r.N1 = j.N1;
    r.m = j.m;
    r.E.resize((size_t)j.m);
    if (j.kind == kDense) {
        j.dense->makeSolution();
        j.dense->get_E(r.E.data());
        r.x0.resize((size_t)j.m * j.N0);
        j.dense->get_x(r.x0.data());
        r.attempts = (long long)j.N0 * j.m * (long long)j.G.size();
        j.dense.reset();
    } else {
        j.bip->makeSolution();
        j.bip->get_E(r.E.data());
        r.x0.resize((size_t)j.m * j.N0);
        r.x1.resize((size_t)j.m * j.N1);
        j.bip->get_x(r.x0.data(), r.x1.data());
        r.attempts = (long long)(j.N0 + j.N1) * j.m * (long long)j.G.size();
        j.bip.reset();
    }
    r.submitS = j.submitS;
    r.admitS = j.admitS;
    r.finishS = now();
    memInUse_ -= j.bytes;
    j.admitted = false;
    done_[j.id] = std::move(r);
    newlyDone_.push_back(j.id);
}

std::vector<int> Server::takeDone() {
    std::vector<int> out;
    out.swap(newlyDone_);
    return out;
}

const Server::Result &Server::result(int id) const {
    auto it = done_.find(id);
    SQ_THROW_IF(it == done_.end(), "job %d has no result (not finished, or released)", id);
    return it->second;
}

void Server::release(int id) { done_.erase(id); }

}  // namespace sqaod_rocm
