/*
 * extern "C" layer of libsqaod_rocm (declarations and reference citations in
 * include/sqaod_rocm.h).  Exceptions never cross the ABI: each entry point
 * catches, records the message for sqaod_rocm_last_error() and returns -1,
 * the reference's TRY/CATCH_ERROR_AND_RETURN pattern (pyglue.h:394-399).
 */
#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>

#include "sqaod_rocm.h"
#include <hip/hip_version.h>

#include "solver.h"

using namespace sqaod_rocm;

namespace {

thread_local std::string g_lastError;

template <class F>
int guarded(F &&f) {
    try {
        f();
        g_lastError.clear();
        return 0;
    } catch (const std::exception &e) {
        g_lastError = e.what();
    } catch (...) {
        g_lastError = "unknown error";
    }
    return -1;
}

struct DgHandle {
    int dtype;
    DenseGraphAnnealer<float> *f = nullptr;
    DenseGraphAnnealer<double> *d = nullptr;
};

template <class Fn>
void withDg(sqaod_rocm_dg_annealer *h, Fn &&fn) {
    SQ_THROW_IF(h == nullptr, "null annealer handle");
    DgHandle *dh = reinterpret_cast<DgHandle *>(h);
    if (dh->dtype == 0)
        fn(*dh->f);
    else
        fn(*dh->d);
}

struct BfHandle {
    int dtype;
    DenseGraphBFSearcher<float> *f = nullptr;
    DenseGraphBFSearcher<double> *d = nullptr;
};

template <class Fn>
void withBf(sqaod_rocm_dg_bf_searcher *h, Fn &&fn) {
    SQ_THROW_IF(h == nullptr, "null searcher handle");
    BfHandle *bh = reinterpret_cast<BfHandle *>(h);
    if (bh->dtype == 0)
        fn(*bh->f);
    else
        fn(*bh->d);
}

struct BgBfHandle {
    int dtype;
    BipartiteGraphBFSearcher<float> *f = nullptr;
    BipartiteGraphBFSearcher<double> *d = nullptr;
};

template <class Fn>
void withBgBf(sqaod_rocm_bg_bf_searcher *h, Fn &&fn) {
    SQ_THROW_IF(h == nullptr, "null searcher handle");
    BgBfHandle *bh = reinterpret_cast<BgBfHandle *>(h);
    if (bh->dtype == 0)
        fn(*bh->f);
    else
        fn(*bh->d);
}

struct FormulasHandle {
    Formulas<float> f;
    Formulas<double> d;
};

template <class Fn>
void withFormulas(sqaod_rocm_formulas *h, int dtype, Fn &&fn) {
    SQ_THROW_IF(h == nullptr, "null formulas handle");
    SQ_THROW_IF(dtype != 0 && dtype != 1, "dtype must be 0 (float32) or 1 (float64)");
    FormulasHandle *fh = reinterpret_cast<FormulasHandle *>(h);
    if (dtype == 0)
        fn(fh->f);
    else
        fn(fh->d);
}

#define SQ_REAL(s) typename std::remove_reference_t<decltype(s)>::real_type

struct BgHandle {
    int dtype;
    BipartiteGraphAnnealer<float> *f = nullptr;
    BipartiteGraphAnnealer<double> *d = nullptr;
};

template <class Fn>
void withBg(sqaod_rocm_bg_annealer *h, Fn &&fn) {
    SQ_THROW_IF(h == nullptr, "null annealer handle");
    BgHandle *bh = reinterpret_cast<BgHandle *>(h);
    if (bh->dtype == 0)
        fn(*bh->f);
    else
        fn(*bh->d);
}

template <class S>
void checkBgShape(S &s, int m, int N0, int N1, const char *what) {
    int n0 = 0, n1 = 0;
    s.getProblemSize(&n0, &n1);
    SQ_THROW_IF(m != s.numTrotters() || N0 != n0 || N1 != n1, "%s: expected m=%d, N0=%d, N1=%d, got %d, %d, %d", what,
                s.numTrotters(), n0, n1, m, N0, N1);
}

std::string joinPrefs(const PrefList &prefs) {
    std::string out;
    for (auto &kv : prefs) {
        if (!out.empty())
            out += ';';
        out += kv.first + "=" + kv.second;
    }
    return out;
}

void copyOut(const std::string &out, char *buf, size_t buflen, size_t *needed) {
    if (needed)
        *needed = out.size() + 1;
    if (buf && buflen) {
        size_t n = out.size() < buflen - 1 ? out.size() : buflen - 1;
        memcpy(buf, out.data(), n);
        buf[n] = '\0';
    }
}

}  // namespace

struct sqaod_rocm_device {
    Device *dev;
};

extern "C" {

const char *sqaod_rocm_last_error(void) { return g_lastError.c_str(); }

void sqaodc_rocm_version(int *version, int *hip_version) {
    if (version)
        *version = 10003; /* the sqaodc solver API this library implements (sqaodc 1.0.3) */
    if (hip_version)
        *hip_version = HIP_VERSION; /* major * 10^7 + minor * 10^5 + patch of the HIP it was built with */
}

int sqaod_rocm_device_count(int *count) {
    return guarded([&] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        *count = (e == hipSuccess) ? n : 0;
    });
}

int sqaod_rocm_device_new(int devno, sqaod_rocm_device **out) {
    return guarded([&] {
        SQ_THROW_IF(out == nullptr, "null output pointer");
        auto *d = new sqaod_rocm_device;
        try {
            d->dev = new Device(devno);
        } catch (...) {
            delete d;
            throw;
        }
        *out = d;
    });
}

int sqaod_rocm_device_delete(sqaod_rocm_device *dev) {
    return guarded([&] {
        if (!dev)
            return;
        delete dev->dev;
        delete dev;
    });
}

int sqaod_rocm_device_synchronize(sqaod_rocm_device *dev) {
    return guarded([&] {
        SQ_THROW_IF(!dev, "null device");
        dev->dev->makeCurrent();
        dev->dev->synchronize();
    });
}

int sqaod_rocm_dg_annealer_new(int dtype, sqaod_rocm_dg_annealer **out) {
    return guarded([&] {
        SQ_THROW_IF(out == nullptr, "null output pointer");
        SQ_THROW_IF(dtype != 0 && dtype != 1, "dtype must be 0 (float32) or 1 (float64)");
        auto *h = new DgHandle;
        h->dtype = dtype;
        if (dtype == 0)
            h->f = new DenseGraphAnnealer<float>();
        else
            h->d = new DenseGraphAnnealer<double>();
        *out = reinterpret_cast<sqaod_rocm_dg_annealer *>(h);
    });
}

int sqaod_rocm_dg_annealer_delete(sqaod_rocm_dg_annealer *a) {
    return guarded([&] {
        if (!a)
            return;
        DgHandle *h = reinterpret_cast<DgHandle *>(a);
        delete h->f;
        delete h->d;
        delete h;
    });
}

int sqaod_rocm_dg_annealer_assign_device(sqaod_rocm_dg_annealer *a, sqaod_rocm_device *dev) {
    return guarded([&] {
        SQ_THROW_IF(!dev, "null device");
        withDg(a, [&](auto &s) { s.assignDevice(*dev->dev); });
    });
}

int sqaod_rocm_dg_annealer_seed(sqaod_rocm_dg_annealer *a, unsigned long long seed) {
    return guarded([&] { withDg(a, [&](auto &s) { s.seed(seed); }); });
}

int sqaod_rocm_dg_annealer_set_qubo(sqaod_rocm_dg_annealer *a, const void *W, int N, int optimize) {
    return guarded([&] {
        SQ_THROW_IF(!W, "null W");
        SQ_THROW_IF(optimize != 0 && optimize != 1, "optimize must be 0 (minimize) or 1 (maximize)");
        SQ_THROW_IF(!a, "null annealer handle");
        DgHandle *h = reinterpret_cast<DgHandle *>(a);
        if (h->dtype == 0)
            h->f->setQUBO((const float *)W, N, (OptimizeMethod)optimize);
        else
            h->d->setQUBO((const double *)W, N, (OptimizeMethod)optimize);
    });
}

int sqaod_rocm_dg_annealer_set_hamiltonian(sqaod_rocm_dg_annealer *a, const void *hv, const void *J, const void *c,
                                           int N) {
    return guarded([&] {
        SQ_THROW_IF(!hv || !J || !c, "null argument");
        SQ_THROW_IF(!a, "null annealer handle");
        DgHandle *h = reinterpret_cast<DgHandle *>(a);
        if (h->dtype == 0)
            h->f->setHamiltonian((const float *)hv, (const float *)J, *(const float *)c, N);
        else
            h->d->setHamiltonian((const double *)hv, (const double *)J, *(const double *)c, N);
    });
}

int sqaod_rocm_dg_annealer_get_problem_size(sqaod_rocm_dg_annealer *a, int *N) {
    return guarded([&] {
        withDg(a, [&](auto &s) {
            s.throwErrorIfProblemNotSet();
            *N = s.getProblemSize();
        });
    });
}

int sqaod_rocm_dg_annealer_set_preference(sqaod_rocm_dg_annealer *a, const char *name, const char *value) {
    return guarded([&] {
        SQ_THROW_IF(!name || !value, "null preference");
        withDg(a, [&](auto &s) { s.setPreference(name, value); });
    });
}

int sqaod_rocm_dg_annealer_get_preferences(sqaod_rocm_dg_annealer *a, char *buf, size_t buflen, size_t *needed) {
    return guarded([&] {
        std::string out;
        withDg(a, [&](auto &s) {
            for (auto &kv : s.getPreferences()) {
                if (!out.empty())
                    out += ';';
                out += kv.first + "=" + kv.second;
            }
        });
        if (needed)
            *needed = out.size() + 1;
        if (buf && buflen) {
            size_t n = out.size() < buflen - 1 ? out.size() : buflen - 1;
            memcpy(buf, out.data(), n);
            buf[n] = 0;
        }
    });
}

int sqaod_rocm_dg_annealer_get_n_trotters(sqaod_rocm_dg_annealer *a, int *m) {
    return guarded([&] { withDg(a, [&](auto &s) { *m = s.numTrotters(); }); });
}

int sqaod_rocm_dg_annealer_get_E(sqaod_rocm_dg_annealer *a, void *E, int m) {
    return guarded([&] {
        SQ_THROW_IF(!a, "null annealer handle");
        DgHandle *h = reinterpret_cast<DgHandle *>(a);
        if (h->dtype == 0) {
            SQ_THROW_IF(m != h->f->numTrotters(), "E buffer has %d entries, %d needed", m, h->f->numTrotters());
            h->f->get_E((float *)E);
        } else {
            SQ_THROW_IF(m != h->d->numTrotters(), "E buffer has %d entries, %d needed", m, h->d->numTrotters());
            h->d->get_E((double *)E);
        }
    });
}

int sqaod_rocm_dg_annealer_get_x(sqaod_rocm_dg_annealer *a, int8_t *x, int m, int N) {
    return guarded([&] {
        withDg(a, [&](auto &s) {
            SQ_THROW_IF(m != s.numTrotters() || N != s.getProblemSize(), "x buffer shape (%d, %d) != (%d, %d)", m, N,
                        s.numTrotters(), s.getProblemSize());
            s.get_x(x);
        });
    });
}

int sqaod_rocm_dg_annealer_get_hamiltonian(sqaod_rocm_dg_annealer *a, void *hv, void *J, void *c) {
    return guarded([&] {
        SQ_THROW_IF(!a, "null annealer handle");
        DgHandle *h = reinterpret_cast<DgHandle *>(a);
        if (h->dtype == 0)
            h->f->getHamiltonian((float *)hv, (float *)J, (float *)c);
        else
            h->d->getHamiltonian((double *)hv, (double *)J, (double *)c);
    });
}

int sqaod_rocm_dg_annealer_get_q(sqaod_rocm_dg_annealer *a, int8_t *q, int m, int N) {
    return guarded([&] {
        withDg(a, [&](auto &s) {
            SQ_THROW_IF(m != s.numTrotters() || N != s.getProblemSize(), "q buffer shape (%d, %d) != (%d, %d)", m, N,
                        s.numTrotters(), s.getProblemSize());
            s.get_q(q);
        });
    });
}

int sqaod_rocm_dg_annealer_set_q(sqaod_rocm_dg_annealer *a, const int8_t *q, int N) {
    return guarded([&] { withDg(a, [&](auto &s) { s.set_q(q, N); }); });
}

int sqaod_rocm_dg_annealer_set_qset(sqaod_rocm_dg_annealer *a, const int8_t *q, int m, int N) {
    return guarded([&] { withDg(a, [&](auto &s) { s.set_qset(q, m, N); }); });
}

int sqaod_rocm_dg_annealer_randomize_spin(sqaod_rocm_dg_annealer *a) {
    return guarded([&] { withDg(a, [&](auto &s) { s.randomizeSpin(); }); });
}

int sqaod_rocm_dg_annealer_calculate_E(sqaod_rocm_dg_annealer *a) {
    return guarded([&] { withDg(a, [&](auto &s) { s.calculate_E(); }); });
}

int sqaod_rocm_dg_annealer_prepare(sqaod_rocm_dg_annealer *a) {
    return guarded([&] { withDg(a, [&](auto &s) { s.prepare(); }); });
}

int sqaod_rocm_dg_annealer_make_solution(sqaod_rocm_dg_annealer *a) {
    return guarded([&] { withDg(a, [&](auto &s) { s.makeSolution(); }); });
}

int sqaod_rocm_dg_annealer_get_system_E(sqaod_rocm_dg_annealer *a, double G, double beta, double *E) {
    return guarded([&] {
        SQ_THROW_IF(!a, "null annealer handle");
        DgHandle *h = reinterpret_cast<DgHandle *>(a);
        if (h->dtype == 0)
            *E = (double)h->f->getSystemE((float)G, (float)beta);
        else
            *E = h->d->getSystemE(G, beta);
    });
}

int sqaod_rocm_dg_annealer_anneal_one_step(sqaod_rocm_dg_annealer *a, double G, double beta) {
    return guarded([&] {
        SQ_THROW_IF(!a, "null annealer handle");
        DgHandle *h = reinterpret_cast<DgHandle *>(a);
        if (h->dtype == 0)
            h->f->annealOneStep((float)G, (float)beta);
        else
            h->d->annealOneStep(G, beta);
    });
}

int sqaod_rocm_dg_annealer_get_counters(sqaod_rocm_dg_annealer *a, unsigned long long counters[3]) {
    return guarded([&] { withDg(a, [&](auto &s) { s.getCounters(counters); }); });
}

int sqaod_rocm_dg_annealer_get_fields(sqaod_rocm_dg_annealer *a, void *F, int m, int N) {
    return guarded([&] {
        SQ_THROW_IF(!a || !F, "null argument");
        DgHandle *h = reinterpret_cast<DgHandle *>(a);
        if (h->dtype == 0) {
            SQ_THROW_IF(m != h->f->numTrotters() || N != h->f->getProblemSize(), "F buffer shape mismatch");
            h->f->getFields((float *)F);
        } else {
            SQ_THROW_IF(m != h->d->numTrotters() || N != h->d->getProblemSize(), "F buffer shape mismatch");
            h->d->getFields((double *)F);
        }
    });
}

int sqaod_rocm_dg_annealer_get_phase_times(sqaod_rocm_dg_annealer *a, double ms[3], unsigned long long counts[3]) {
    return guarded([&] { withDg(a, [&](auto &s) { s.getPhaseTimes(ms, counts); }); });
}

int sqaod_rocm_dg_annealer_get_energy_rounding_error(sqaod_rocm_dg_annealer *a, double *err) {
    return guarded([&] { withDg(a, [&](auto &s) { *err = s.lastEnergyRoundingError(); }); });
}

int sqaod_rocm_dg_annealer_get_E_fp64(sqaod_rocm_dg_annealer *a, double *E, int m) {
    return guarded([&] {
        withDg(a, [&](auto &s) {
            const std::vector<double> &e = s.energiesFP64();
            SQ_THROW_IF(e.empty(), "get_E_fp64: no energies calculated yet");
            SQ_THROW_IF(m != (int)e.size() || !E, "get_E_fp64: expected %d entries, got %d", (int)e.size(), m);
            std::copy(e.begin(), e.end(), E);
        });
    });
}

int sqaod_rocm_dg_annealer_set_system_E_trace(sqaod_rocm_dg_annealer *a, int capacity) {
    return guarded([&] { withDg(a, [&](auto &s) { s.setSystemETrace(capacity); }); });
}

int sqaod_rocm_dg_annealer_get_system_E_trace(sqaod_rocm_dg_annealer *a, double *out, int cap, int *n,
                                              long long *steps) {
    return guarded([&] { withDg(a, [&](auto &s) { *n = s.getSystemETrace(out, cap, steps); }); });
}

int sqaod_rocm_dg_annealer_get_form_energy_error(sqaod_rocm_dg_annealer *a, double *err) {
    return guarded([&] { withDg(a, [&](auto &s) { *err = s.formEnergyError(); }); });
}

int sqaod_rocm_expf_checksum(unsigned long long lo, unsigned long long hi, unsigned long long *out) {
    return guarded([&] {
        SQ_THROW_IF(!out || lo > hi || hi > 0x100000000ull, "bad range");
        defaultDevice().makeCurrent();
        *out = expfChecksum(lo, hi);
    });
}

/* ---------------- dense batch ---------------- */

extern "C++" { /* internal helpers returning C++ references */
namespace {
DenseBatch &batchOf(sqaod_rocm_dense_batch *b) {
    SQ_THROW_IF(b == nullptr, "null batch handle");
    return *reinterpret_cast<DenseBatch *>(b);
}
DenseGraphAnnealer<float> &batchInstance(sqaod_rocm_dense_batch *b, int i, int m, int N) {
    auto &a = batchOf(b).instance(i);
    SQ_THROW_IF(m != a.numTrotters() || (N >= 0 && N != a.getProblemSize()), "buffer shape (%d, %d) != (%d, %d)", m,
                N, a.numTrotters(), a.getProblemSize());
    return a;
}
}  // namespace
}  // extern "C++"

int sqaod_rocm_dense_batch_new(sqaod_rocm_dense_batch **out) {
    return guarded([&] {
        SQ_THROW_IF(!out, "null output");
        *out = reinterpret_cast<sqaod_rocm_dense_batch *>(new DenseBatch(defaultDevice()));
    });
}

int sqaod_rocm_dense_batch_delete(sqaod_rocm_dense_batch *b) {
    return guarded([&] { delete &batchOf(b); });
}

int sqaod_rocm_dense_batch_set_preference(sqaod_rocm_dense_batch *b, const char *name, const char *value) {
    return guarded([&] {
        SQ_THROW_IF(!name || !value, "null preference");
        batchOf(b).setPreference(name, value);
    });
}

int sqaod_rocm_dense_batch_add_qubo(sqaod_rocm_dense_batch *b, const float *W, int N, int m, unsigned long long seed,
                                    int maximize, int *index) {
    return guarded([&] {
        SQ_THROW_IF(!W || !index, "null argument");
        *index = batchOf(b).addQUBO(W, N, m, seed, maximize ? optMaximize : optMinimize);
    });
}

int sqaod_rocm_dense_batch_size(sqaod_rocm_dense_batch *b, int *n) {
    return guarded([&] { *n = batchOf(b).size(); });
}

int sqaod_rocm_dense_batch_prepare(sqaod_rocm_dense_batch *b) {
    return guarded([&] { batchOf(b).prepare(); });
}

int sqaod_rocm_dense_batch_anneal(sqaod_rocm_dense_batch *b, const double *G, const double *beta, int nSteps,
                                  int graph, double *ms) {
    return guarded([&] {
        SQ_THROW_IF(!G || !beta, "null schedule");
        batchOf(b).anneal(G, beta, nSteps, graph != 0);
        if (ms)
            *ms = batchOf(b).lastAnnealMs();
    });
}

int sqaod_rocm_dense_batch_autotune(sqaod_rocm_dense_batch *b, double G, double beta, int trialSteps,
                                    int *wavesPerBlock, int *largestFirst) {
    return guarded([&] {
        batchOf(b).autotune(G, beta, trialSteps);
        if (wavesPerBlock)
            *wavesPerBlock = batchOf(b).wavesPerBlock();
        if (largestFirst)
            *largestFirst = batchOf(b).largestFirst() ? 1 : 0;
    });
}

int sqaod_rocm_dense_batch_attempts_per_step(sqaod_rocm_dense_batch *b, long long *n) {
    return guarded([&] { *n = batchOf(b).attemptsPerStep(); });
}

int sqaod_rocm_dense_batch_get_shape(sqaod_rocm_dense_batch *b, int i, int *N, int *m) {
    return guarded([&] {
        auto &a = batchOf(b).instance(i);
        *N = a.getProblemSize();
        *m = a.numTrotters();
    });
}

int sqaod_rocm_dense_batch_get_E(sqaod_rocm_dense_batch *b, int i, float *E, int m) {
    return guarded([&] { batchInstance(b, i, m, -1).get_E(E); });
}

int sqaod_rocm_dense_batch_get_x(sqaod_rocm_dense_batch *b, int i, int8_t *x, int m, int N) {
    return guarded([&] { batchInstance(b, i, m, N).get_x(x); });
}

int sqaod_rocm_dense_batch_get_q(sqaod_rocm_dense_batch *b, int i, int8_t *q, int m, int N) {
    return guarded([&] { batchInstance(b, i, m, N).get_q(q); });
}

/* ---------------- serving ---------------- */

extern "C++" { /* internal helper returning a C++ reference */
namespace {
Server &serverOf(sqaod_rocm_server *s) {
    SQ_THROW_IF(s == nullptr, "null server handle");
    return *reinterpret_cast<Server *>(s);
}
}  // namespace
}  // extern "C++"

int sqaod_rocm_server_new(sqaod_rocm_server **out) {
    return guarded([&] {
        SQ_THROW_IF(!out, "null output");
        *out = reinterpret_cast<sqaod_rocm_server *>(new Server(defaultDevice()));
    });
}

int sqaod_rocm_server_delete(sqaod_rocm_server *s) {
    return guarded([&] { delete &serverOf(s); });
}

int sqaod_rocm_server_set_preference(sqaod_rocm_server *s, const char *name, const char *value) {
    return guarded([&] {
        SQ_THROW_IF(!name || !value, "null preference");
        serverOf(s).setPreference(name, value);
    });
}

int sqaod_rocm_server_submit_dense(sqaod_rocm_server *s, const float *W, int N, int m, unsigned long long seed,
                                   const double *G, const double *beta, int nSteps, int *id) {
    return guarded([&] {
        SQ_THROW_IF(!id, "null id");
        *id = serverOf(s).submitDense(W, N, m, seed, G, beta, nSteps);
    });
}

int sqaod_rocm_server_submit_bipartite(sqaod_rocm_server *s, const float *b0, const float *b1, const float *W, int N0,
                                       int N1, int m, unsigned long long seed, const double *G, const double *beta,
                                       int nSteps, int *id) {
    return guarded([&] {
        SQ_THROW_IF(!id, "null id");
        *id = serverOf(s).submitBipartite(b0, b1, W, N0, N1, m, seed, G, beta, nSteps);
    });
}

int sqaod_rocm_server_tick(sqaod_rocm_server *s, int *remaining) {
    return guarded([&] {
        const int r = serverOf(s).tick();
        if (remaining)
            *remaining = r;
    });
}

int sqaod_rocm_server_status(sqaod_rocm_server *s, int *pending, int *active, unsigned long long *bytesInUse,
                             unsigned long long *peakBytes) {
    return guarded([&] {
        Server &sv = serverOf(s);
        if (pending)
            *pending = sv.numPending();
        if (active)
            *active = sv.numActive();
        if (bytesInUse)
            *bytesInUse = sv.memoryInUse();
        if (peakBytes)
            *peakBytes = sv.peakMemory();
    });
}

int sqaod_rocm_server_take_done(sqaod_rocm_server *s, int *ids, int cap, int *n) {
    return guarded([&] {
        SQ_THROW_IF(!ids || !n || cap < 0, "bad output");
        std::vector<int> d = serverOf(s).takeDone();
        SQ_THROW_IF((int)d.size() > cap, "%zu finished jobs, room for %d", d.size(), cap);
        std::copy(d.begin(), d.end(), ids);
        *n = (int)d.size();
    });
}

int sqaod_rocm_server_result_shape(sqaod_rocm_server *s, int id, int *kind, int *N0, int *N1, int *m) {
    return guarded([&] {
        const Server::Result &r = serverOf(s).result(id);
        *kind = r.kind;
        *N0 = r.N0;
        *N1 = r.N1;
        *m = r.m;
    });
}

int sqaod_rocm_server_result(sqaod_rocm_server *s, int id, float *E, int8_t *x0, int8_t *x1, double times[3],
                             long long *attempts) {
    return guarded([&] {
        const Server::Result &r = serverOf(s).result(id);
        SQ_THROW_IF(!r.error.empty(), "job %d failed: %s", id, r.error.c_str());
        if (E)
            std::copy(r.E.begin(), r.E.end(), E);
        if (x0)
            std::copy(r.x0.begin(), r.x0.end(), x0);
        if (x1)
            std::copy(r.x1.begin(), r.x1.end(), x1);
        if (times) {
            times[0] = r.submitS;
            times[1] = r.admitS;
            times[2] = r.finishS;
        }
        if (attempts)
            *attempts = r.attempts;
    });
}

int sqaod_rocm_server_result_error(sqaod_rocm_server *s, int id, char *buf, size_t buflen, int *failed) {
    return guarded([&] {
        const Server::Result &r = serverOf(s).result(id);
        *failed = r.error.empty() ? 0 : 1;
        if (buf && buflen) {
            const size_t n = std::min(buflen - 1, r.error.size());
            memcpy(buf, r.error.data(), n);
            buf[n] = 0;
        }
    });
}

int sqaod_rocm_server_release(sqaod_rocm_server *s, int id) {
    return guarded([&] { serverOf(s).release(id); });
}

int sqaod_rocm_server_now(sqaod_rocm_server *s, double *t) {
    return guarded([&] { *t = serverOf(s).now(); });
}

int sqaod_rocm_fp8_decode_table(float *out) {
    return guarded([&] {
        SQ_THROW_IF(!out, "null output");
        defaultDevice().makeCurrent();
        fp8DecodeTable(out);
    });
}

/* ---------------- dense graph brute-force searcher ---------------- */

int sqaod_rocm_dg_bf_searcher_new(int dtype, sqaod_rocm_dg_bf_searcher **out) {
    return guarded([&] {
        SQ_THROW_IF(out == nullptr, "null output pointer");
        SQ_THROW_IF(dtype != 0 && dtype != 1, "dtype must be 0 (float32) or 1 (float64)");
        auto *h = new BfHandle;
        h->dtype = dtype;
        if (dtype == 0)
            h->f = new DenseGraphBFSearcher<float>();
        else
            h->d = new DenseGraphBFSearcher<double>();
        *out = reinterpret_cast<sqaod_rocm_dg_bf_searcher *>(h);
    });
}

int sqaod_rocm_dg_bf_searcher_delete(sqaod_rocm_dg_bf_searcher *b) {
    return guarded([&] {
        if (!b)
            return;
        BfHandle *h = reinterpret_cast<BfHandle *>(b);
        delete h->f;
        delete h->d;
        delete h;
    });
}

int sqaod_rocm_dg_bf_searcher_assign_device(sqaod_rocm_dg_bf_searcher *b, sqaod_rocm_device *dev) {
    return guarded([&] {
        SQ_THROW_IF(!dev, "null device");
        withBf(b, [&](auto &s) { s.assignDevice(*dev->dev); });
    });
}

int sqaod_rocm_dg_bf_searcher_set_qubo(sqaod_rocm_dg_bf_searcher *b, const void *W, int N, int optimize) {
    return guarded([&] {
        SQ_THROW_IF(!W, "null W");
        withBf(b, [&](auto &s) {
            using real = typename std::remove_reference_t<decltype(s)>::real_type;
            s.setQUBO(static_cast<const real *>(W), N, optimize ? optMaximize : optMinimize);
        });
    });
}

int sqaod_rocm_dg_bf_searcher_get_problem_size(sqaod_rocm_dg_bf_searcher *b, int *N) {
    return guarded([&] {
        withBf(b, [&](auto &s) {
            s.throwErrorIfProblemNotSet();
            *N = s.getProblemSize();
        });
    });
}

int sqaod_rocm_dg_bf_searcher_set_preference(sqaod_rocm_dg_bf_searcher *b, const char *name, const char *value) {
    return guarded([&] {
        SQ_THROW_IF(!name || !value, "null preference");
        withBf(b, [&](auto &s) { s.setPreference(name, value); });
    });
}

int sqaod_rocm_dg_bf_searcher_get_preferences(sqaod_rocm_dg_bf_searcher *b, char *buf, size_t buflen,
                                              size_t *needed) {
    return guarded([&] {
        std::string out;
        withBf(b, [&](auto &s) { out = joinPrefs(s.getPreferences()); });
        copyOut(out, buf, buflen, needed);
    });
}

int sqaod_rocm_dg_bf_searcher_prepare(sqaod_rocm_dg_bf_searcher *b) {
    return guarded([&] { withBf(b, [&](auto &s) { s.prepare(); }); });
}

int sqaod_rocm_dg_bf_searcher_calculate_E(sqaod_rocm_dg_bf_searcher *b) {
    return guarded([&] { withBf(b, [&](auto &s) { s.calculate_E(); }); });
}

int sqaod_rocm_dg_bf_searcher_make_solution(sqaod_rocm_dg_bf_searcher *b) {
    return guarded([&] { withBf(b, [&](auto &s) { s.makeSolution(); }); });
}

int sqaod_rocm_dg_bf_searcher_search_range(sqaod_rocm_dg_bf_searcher *b, int *done, unsigned long long *cur_x_end) {
    return guarded([&] {
        withBf(b, [&](auto &s) {
            unsigned long long x = 0;
            bool d = s.searchRange(&x);
            if (done)
                *done = d ? 1 : 0;
            if (cur_x_end)
                *cur_x_end = x;
        });
    });
}

int sqaod_rocm_dg_bf_searcher_search(sqaod_rocm_dg_bf_searcher *b) {
    return guarded([&] { withBf(b, [&](auto &s) { s.search(); }); });
}

int sqaod_rocm_dense_bf_search_noalloc(int dtype, const void *d_W, int N, int optimize, unsigned long long tile_size,
                                       void *d_Emin, unsigned long long *d_x, unsigned long long *d_n, void *stream) {
    return guarded(
        [&] { denseBFSearchNoalloc(dtype, d_W, N, optimize, tile_size, d_Emin, d_x, d_n, (hipStream_t)stream); });
}

int sqaod_rocm_dg_bf_searcher_get_n_solutions(sqaod_rocm_dg_bf_searcher *b, int *n) {
    return guarded([&] {
        SQ_THROW_IF(!n, "null output pointer");
        withBf(b, [&](auto &s) { *n = s.numSolutions(); });
    });
}

int sqaod_rocm_dg_bf_searcher_get_x(sqaod_rocm_dg_bf_searcher *b, int8_t *x, int n, int N) {
    return guarded([&] {
        withBf(b, [&](auto &s) {
            const int ns = s.numSolutions();
            SQ_THROW_IF(n != ns || N != s.getProblemSize(), "get_x: expected %d x %d, got %d x %d", ns,
                        s.getProblemSize(), n, N);
            if (ns)
                s.get_x(x);
        });
    });
}

int sqaod_rocm_dg_bf_searcher_get_E(sqaod_rocm_dg_bf_searcher *b, void *E, int n) {
    return guarded([&] {
        withBf(b, [&](auto &s) {
            using real = typename std::remove_reference_t<decltype(s)>::real_type;
            const int ne = s.numEnergies();
            SQ_THROW_IF(n != ne, "get_E: expected %d entries, got %d", ne, n);
            s.get_E(static_cast<real *>(E));
        });
    });
}

/* ---------------- bipartite graph brute-force searcher ---------------- */

int sqaod_rocm_bg_bf_searcher_new(int dtype, sqaod_rocm_bg_bf_searcher **out) {
    return guarded([&] {
        SQ_THROW_IF(out == nullptr, "null output pointer");
        SQ_THROW_IF(dtype != 0 && dtype != 1, "dtype must be 0 (float32) or 1 (float64)");
        auto *h = new BgBfHandle;
        h->dtype = dtype;
        if (dtype == 0)
            h->f = new BipartiteGraphBFSearcher<float>();
        else
            h->d = new BipartiteGraphBFSearcher<double>();
        *out = reinterpret_cast<sqaod_rocm_bg_bf_searcher *>(h);
    });
}

int sqaod_rocm_bg_bf_searcher_delete(sqaod_rocm_bg_bf_searcher *b) {
    return guarded([&] {
        if (!b)
            return;
        BgBfHandle *h = reinterpret_cast<BgBfHandle *>(b);
        delete h->f;
        delete h->d;
        delete h;
    });
}

int sqaod_rocm_bg_bf_searcher_assign_device(sqaod_rocm_bg_bf_searcher *b, sqaod_rocm_device *dev) {
    return guarded([&] {
        SQ_THROW_IF(!dev, "null device");
        withBgBf(b, [&](auto &s) { s.assignDevice(*dev->dev); });
    });
}

int sqaod_rocm_bg_bf_searcher_set_qubo(sqaod_rocm_bg_bf_searcher *b, const void *b0, int N0, const void *b1, int N1,
                                       const void *W, int optimize) {
    return guarded([&] {
        SQ_THROW_IF(!b0 || !b1 || !W, "null b0, b1 or W");
        withBgBf(b, [&](auto &s) {
            using real = typename std::remove_reference_t<decltype(s)>::real_type;
            s.setQUBO(static_cast<const real *>(b0), static_cast<const real *>(b1), static_cast<const real *>(W), N0,
                      N1, optimize ? optMaximize : optMinimize);
        });
    });
}

int sqaod_rocm_bg_bf_searcher_get_problem_size(sqaod_rocm_bg_bf_searcher *b, int *N0, int *N1) {
    return guarded([&] { withBgBf(b, [&](auto &s) { s.getProblemSize(N0, N1); }); });
}

int sqaod_rocm_bg_bf_searcher_set_preference(sqaod_rocm_bg_bf_searcher *b, const char *name, const char *value) {
    return guarded([&] {
        SQ_THROW_IF(!name || !value, "null preference");
        withBgBf(b, [&](auto &s) { s.setPreference(name, value); });
    });
}

int sqaod_rocm_bg_bf_searcher_get_preferences(sqaod_rocm_bg_bf_searcher *b, char *buf, size_t buflen,
                                              size_t *needed) {
    return guarded([&] {
        std::string out;
        withBgBf(b, [&](auto &s) { out = joinPrefs(s.getPreferences()); });
        copyOut(out, buf, buflen, needed);
    });
}

int sqaod_rocm_bg_bf_searcher_prepare(sqaod_rocm_bg_bf_searcher *b) {
    return guarded([&] { withBgBf(b, [&](auto &s) { s.prepare(); }); });
}

int sqaod_rocm_bg_bf_searcher_calculate_E(sqaod_rocm_bg_bf_searcher *b) {
    return guarded([&] { withBgBf(b, [&](auto &s) { s.calculate_E(); }); });
}

int sqaod_rocm_bg_bf_searcher_make_solution(sqaod_rocm_bg_bf_searcher *b) {
    return guarded([&] { withBgBf(b, [&](auto &s) { s.makeSolution(); }); });
}

int sqaod_rocm_bg_bf_searcher_search_range(sqaod_rocm_bg_bf_searcher *b, int *done, unsigned long long *cur_x0,
                                           unsigned long long *cur_x1) {
    return guarded([&] {
        withBgBf(b, [&](auto &s) {
            unsigned long long x0 = 0, x1 = 0;
            bool d = s.searchRange(&x0, &x1);
            if (done)
                *done = d ? 1 : 0;
            if (cur_x0)
                *cur_x0 = x0;
            if (cur_x1)
                *cur_x1 = x1;
        });
    });
}

int sqaod_rocm_bg_bf_searcher_search(sqaod_rocm_bg_bf_searcher *b) {
    return guarded([&] { withBgBf(b, [&](auto &s) { s.search(); }); });
}

int sqaod_rocm_bg_bf_searcher_get_n_solutions(sqaod_rocm_bg_bf_searcher *b, int *n) {
    return guarded([&] {
        SQ_THROW_IF(!n, "null output pointer");
        withBgBf(b, [&](auto &s) { *n = s.numSolutions(); });
    });
}

int sqaod_rocm_bg_bf_searcher_get_x(sqaod_rocm_bg_bf_searcher *b, int8_t *x0, int8_t *x1, int n, int N0, int N1) {
    return guarded([&] {
        withBgBf(b, [&](auto &s) {
            const int ns = s.numSolutions();
            int P0 = 0, P1 = 0;
            s.getProblemSize(&P0, &P1);
            SQ_THROW_IF(n != ns || N0 != P0 || N1 != P1, "get_x: expected %d pairs of (%d, %d), got %d of (%d, %d)",
                        ns, P0, P1, n, N0, N1);
            if (ns)
                s.get_x(x0, x1);
        });
    });
}

int sqaod_rocm_bg_bf_searcher_get_E(sqaod_rocm_bg_bf_searcher *b, void *E, int n) {
    return guarded([&] {
        withBgBf(b, [&](auto &s) {
            using real = typename std::remove_reference_t<decltype(s)>::real_type;
            const int ne = s.numEnergies();
            SQ_THROW_IF(n != ne, "get_E: expected %d entries, got %d", ne, n);
            s.get_E(static_cast<real *>(E));
        });
    });
}

/* ---------------- bipartite graph annealer ---------------- */

int sqaod_rocm_bg_annealer_new(int dtype, sqaod_rocm_bg_annealer **out) {
    return guarded([&] {
        SQ_THROW_IF(out == nullptr, "null output pointer");
        SQ_THROW_IF(dtype != 0 && dtype != 1, "dtype must be 0 (float32) or 1 (float64)");
        auto *h = new BgHandle;
        h->dtype = dtype;
        if (dtype == 0)
            h->f = new BipartiteGraphAnnealer<float>();
        else
            h->d = new BipartiteGraphAnnealer<double>();
        *out = reinterpret_cast<sqaod_rocm_bg_annealer *>(h);
    });
}

int sqaod_rocm_bg_annealer_delete(sqaod_rocm_bg_annealer *a) {
    return guarded([&] {
        if (!a)
            return;
        BgHandle *h = reinterpret_cast<BgHandle *>(a);
        delete h->f;
        delete h->d;
        delete h;
    });
}

int sqaod_rocm_bg_annealer_assign_device(sqaod_rocm_bg_annealer *a, sqaod_rocm_device *dev) {
    return guarded([&] {
        SQ_THROW_IF(!dev, "null device");
        withBg(a, [&](auto &s) { s.assignDevice(*dev->dev); });
    });
}

int sqaod_rocm_bg_annealer_seed(sqaod_rocm_bg_annealer *a, unsigned long long seed) {
    return guarded([&] { withBg(a, [&](auto &s) { s.seed(seed); }); });
}

int sqaod_rocm_bg_annealer_set_qubo(sqaod_rocm_bg_annealer *a, const void *b0, int N0, const void *b1, int N1,
                                    const void *W, int optimize) {
    return guarded([&] {
        SQ_THROW_IF(!b0 || !b1 || !W, "null input");
        withBg(a, [&](auto &s) {
            using real = typename std::remove_reference_t<decltype(s)>::real_type;
            s.setQUBO(static_cast<const real *>(b0), static_cast<const real *>(b1), static_cast<const real *>(W), N0,
                      N1, optimize ? optMaximize : optMinimize);
        });
    });
}

int sqaod_rocm_bg_annealer_set_hamiltonian(sqaod_rocm_bg_annealer *a, const void *h0, int N0, const void *h1, int N1,
                                           const void *J, const void *c) {
    return guarded([&] {
        SQ_THROW_IF(!h0 || !h1 || !J || !c, "null input");
        withBg(a, [&](auto &s) {
            using real = typename std::remove_reference_t<decltype(s)>::real_type;
            s.setHamiltonian(static_cast<const real *>(h0), static_cast<const real *>(h1),
                             static_cast<const real *>(J), *static_cast<const real *>(c), N0, N1);
        });
    });
}

int sqaod_rocm_bg_annealer_get_problem_size(sqaod_rocm_bg_annealer *a, int *N0, int *N1) {
    return guarded([&] {
        SQ_THROW_IF(!N0 || !N1, "null output pointer");
        withBg(a, [&](auto &s) { s.getProblemSize(N0, N1); });
    });
}

int sqaod_rocm_bg_annealer_set_preference(sqaod_rocm_bg_annealer *a, const char *name, const char *value) {
    return guarded([&] {
        SQ_THROW_IF(!name || !value, "null preference");
        withBg(a, [&](auto &s) { s.setPreference(name, value); });
    });
}

int sqaod_rocm_bg_annealer_get_preferences(sqaod_rocm_bg_annealer *a, char *buf, size_t buflen, size_t *needed) {
    return guarded([&] {
        std::string out;
        withBg(a, [&](auto &s) { out = joinPrefs(s.getPreferences()); });
        copyOut(out, buf, buflen, needed);
    });
}

int sqaod_rocm_bg_annealer_get_n_trotters(sqaod_rocm_bg_annealer *a, int *m) {
    return guarded([&] {
        SQ_THROW_IF(!m, "null output pointer");
        withBg(a, [&](auto &s) { *m = s.numTrotters(); });
    });
}

int sqaod_rocm_bg_annealer_get_E(sqaod_rocm_bg_annealer *a, void *E, int m) {
    return guarded([&] {
        withBg(a, [&](auto &s) {
            using real = typename std::remove_reference_t<decltype(s)>::real_type;
            SQ_THROW_IF(m != s.numTrotters(), "get_E: expected %d entries, got %d", s.numTrotters(), m);
            s.get_E(static_cast<real *>(E));
        });
    });
}

int sqaod_rocm_bg_annealer_get_E_fp64(sqaod_rocm_bg_annealer *a, double *E, int m) {
    return guarded([&] {
        withBg(a, [&](auto &s) {
            const std::vector<double> &e = s.energiesFP64();
            SQ_THROW_IF(e.empty(), "get_E_fp64: no energies calculated yet");
            SQ_THROW_IF(m != (int)e.size() || !E, "get_E_fp64: expected %d entries, got %d", (int)e.size(), m);
            std::copy(e.begin(), e.end(), E);
        });
    });
}

int sqaod_rocm_bg_annealer_get_x(sqaod_rocm_bg_annealer *a, int8_t *x0, int8_t *x1, int m, int N0, int N1) {
    return guarded([&] {
        withBg(a, [&](auto &s) {
            checkBgShape(s, m, N0, N1, "get_x");
            s.get_x(x0, x1);
        });
    });
}

int sqaod_rocm_bg_annealer_get_hamiltonian(sqaod_rocm_bg_annealer *a, void *h0, void *h1, void *J, void *c) {
    return guarded([&] {
        withBg(a, [&](auto &s) {
            using real = typename std::remove_reference_t<decltype(s)>::real_type;
            s.getHamiltonian(static_cast<real *>(h0), static_cast<real *>(h1), static_cast<real *>(J),
                             static_cast<real *>(c));
        });
    });
}

int sqaod_rocm_bg_annealer_get_q(sqaod_rocm_bg_annealer *a, int8_t *q0, int8_t *q1, int m, int N0, int N1) {
    return guarded([&] {
        withBg(a, [&](auto &s) {
            checkBgShape(s, m, N0, N1, "get_q");
            s.get_q(q0, q1);
        });
    });
}

int sqaod_rocm_bg_annealer_set_q(sqaod_rocm_bg_annealer *a, const int8_t *q0, int N0, const int8_t *q1, int N1) {
    return guarded([&] { withBg(a, [&](auto &s) { s.set_q(q0, N0, q1, N1); }); });
}

int sqaod_rocm_bg_annealer_set_qset(sqaod_rocm_bg_annealer *a, const int8_t *q0, const int8_t *q1, int m, int N0,
                                    int N1) {
    return guarded([&] { withBg(a, [&](auto &s) { s.set_qset(q0, q1, m, N0, N1); }); });
}

int sqaod_rocm_bg_annealer_randomize_spin(sqaod_rocm_bg_annealer *a) {
    return guarded([&] { withBg(a, [&](auto &s) { s.randomizeSpin(); }); });
}

int sqaod_rocm_bg_annealer_calculate_E(sqaod_rocm_bg_annealer *a) {
    return guarded([&] { withBg(a, [&](auto &s) { s.calculate_E(); }); });
}

int sqaod_rocm_bg_annealer_prepare(sqaod_rocm_bg_annealer *a) {
    return guarded([&] { withBg(a, [&](auto &s) { s.prepare(); }); });
}

int sqaod_rocm_bg_annealer_make_solution(sqaod_rocm_bg_annealer *a) {
    return guarded([&] { withBg(a, [&](auto &s) { s.makeSolution(); }); });
}

int sqaod_rocm_bg_annealer_get_system_E(sqaod_rocm_bg_annealer *a, double G, double beta, double *E) {
    return guarded([&] {
        SQ_THROW_IF(!E, "null output pointer");
        withBg(a, [&](auto &s) {
            using real = typename std::remove_reference_t<decltype(s)>::real_type;
            *E = (double)s.getSystemE((real)G, (real)beta);
        });
    });
}

int sqaod_rocm_bg_annealer_anneal_one_step(sqaod_rocm_bg_annealer *a, double G, double beta) {
    return guarded([&] {
        withBg(a, [&](auto &s) {
            using real = typename std::remove_reference_t<decltype(s)>::real_type;
            s.annealOneStep((real)G, (real)beta);
        });
    });
}

/* ---------------- formulas ---------------- */

int sqaod_rocm_formulas_new(sqaod_rocm_formulas **out) {
    return guarded([&] {
        SQ_THROW_IF(out == nullptr, "null output pointer");
        *out = reinterpret_cast<sqaod_rocm_formulas *>(new FormulasHandle);
    });
}

int sqaod_rocm_formulas_delete(sqaod_rocm_formulas *f) {
    return guarded([&] { delete reinterpret_cast<FormulasHandle *>(f); });
}

int sqaod_rocm_formulas_assign_device(sqaod_rocm_formulas *f, sqaod_rocm_device *dev) {
    return guarded([&] {
        SQ_THROW_IF(!f || !dev, "null handle");
        FormulasHandle *fh = reinterpret_cast<FormulasHandle *>(f);
        fh->f.assignDevice(*dev->dev);
        fh->d.assignDevice(*dev->dev);
    });
}

int sqaod_rocm_dense_graph_batch_calculate_E(sqaod_rocm_formulas *f, int dtype, void *E, const void *W, int N,
                                             const int8_t *x, int nBatch) {
    return guarded([&] {
        SQ_THROW_IF(!E || !W || !x, "null input");
        withFormulas(f, dtype, [&](auto &s) {
            using real = SQ_REAL(s);
            s.denseCalculateE(static_cast<real *>(E), static_cast<const real *>(W), N, x, nBatch);
        });
    });
}

int sqaod_rocm_dense_graph_calculate_hamiltonian(sqaod_rocm_formulas *f, int dtype, void *h, void *J, void *c,
                                                 const void *W, int N) {
    return guarded([&] {
        SQ_THROW_IF(!h || !J || !c || !W, "null input");
        withFormulas(f, dtype, [&](auto &s) {
            using real = SQ_REAL(s);
            s.denseHamiltonian(static_cast<real *>(h), static_cast<real *>(J), static_cast<real *>(c),
                               static_cast<const real *>(W), N);
        });
    });
}

int sqaod_rocm_dense_graph_batch_calculate_E_from_spin(sqaod_rocm_formulas *f, int dtype, void *E, const void *h,
                                                       const void *J, const void *c, const int8_t *q, int N,
                                                       int nBatch) {
    return guarded([&] {
        SQ_THROW_IF(!E || !h || !J || !c || !q, "null input");
        withFormulas(f, dtype, [&](auto &s) {
            using real = SQ_REAL(s);
            s.denseCalculateEFromSpin(static_cast<real *>(E), static_cast<const real *>(h),
                                      static_cast<const real *>(J), *static_cast<const real *>(c), q, N, nBatch);
        });
    });
}

int sqaod_rocm_bipartite_graph_batch_calculate_E(sqaod_rocm_formulas *f, int dtype, void *E, const void *b0, int N0,
                                                 const void *b1, int N1, const void *W, const int8_t *x0,
                                                 const int8_t *x1, int nBatch) {
    return guarded([&] {
        SQ_THROW_IF(!E || !b0 || !b1 || !W || !x0 || !x1, "null input");
        withFormulas(f, dtype, [&](auto &s) {
            using real = SQ_REAL(s);
            s.bipartiteCalculateE(static_cast<real *>(E), static_cast<const real *>(b0), static_cast<const real *>(b1),
                                  static_cast<const real *>(W), N0, N1, x0, nBatch, x1, nBatch, false);
        });
    });
}

int sqaod_rocm_bipartite_graph_batch_calculate_E_2d(sqaod_rocm_formulas *f, int dtype, void *E, const void *b0,
                                                    int N0, const void *b1, int N1, const void *W, const int8_t *x0,
                                                    int nBatch0, const int8_t *x1, int nBatch1) {
    return guarded([&] {
        SQ_THROW_IF(!E || !b0 || !b1 || !W || !x0 || !x1, "null input");
        withFormulas(f, dtype, [&](auto &s) {
            using real = SQ_REAL(s);
            s.bipartiteCalculateE(static_cast<real *>(E), static_cast<const real *>(b0), static_cast<const real *>(b1),
                                  static_cast<const real *>(W), N0, N1, x0, nBatch0, x1, nBatch1, true);
        });
    });
}

int sqaod_rocm_bipartite_graph_calculate_hamiltonian(sqaod_rocm_formulas *f, int dtype, void *h0, void *h1, void *J,
                                                     void *c, const void *b0, int N0, const void *b1, int N1,
                                                     const void *W) {
    return guarded([&] {
        SQ_THROW_IF(!h0 || !h1 || !J || !c || !b0 || !b1 || !W, "null input");
        withFormulas(f, dtype, [&](auto &s) {
            using real = SQ_REAL(s);
            s.bipartiteHamiltonian(static_cast<real *>(h0), static_cast<real *>(h1), static_cast<real *>(J),
                                   static_cast<real *>(c), static_cast<const real *>(b0), static_cast<const real *>(b1),
                                   static_cast<const real *>(W), N0, N1);
        });
    });
}

int sqaod_rocm_bipartite_graph_batch_calculate_E_from_spin(sqaod_rocm_formulas *f, int dtype, void *E, const void *h0,
                                                           int N0, const void *h1, int N1, const void *J,
                                                           const void *c, const int8_t *q0, const int8_t *q1,
                                                           int nBatch) {
    return guarded([&] {
        SQ_THROW_IF(!E || !h0 || !h1 || !J || !c || !q0 || !q1, "null input");
        withFormulas(f, dtype, [&](auto &s) {
            using real = SQ_REAL(s);
            s.bipartiteCalculateEFromSpin(static_cast<real *>(E), static_cast<const real *>(h0),
                                          static_cast<const real *>(h1), static_cast<const real *>(J),
                                          *static_cast<const real *>(c), N0, N1, q0, q1, nBatch);
        });
    });
}

}  // extern "C"
