/*
 * extern "C" layer of libsqaod_rocm (declarations and reference citations in
 * include/sqaod_rocm.h).  Exceptions never cross the ABI: each entry point
 * catches, records the message for sqaod_rocm_last_error() and returns -1,
 * the reference's TRY/CATCH_ERROR_AND_RETURN pattern (pyglue.h:394-399).
 */
#include <cstring>
#include <string>

#include "sqaod_rocm.h"
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

}  // namespace

struct sqaod_rocm_device {
    Device *dev;
};

extern "C" {

const char *sqaod_rocm_last_error(void) { return g_lastError.c_str(); }

void sqaod_rocm_version(int *major, int *minor) {
    if (major)
        *major = 1;
    if (minor)
        *minor = 0;
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

int sqaod_rocm_expf_checksum(unsigned long long lo, unsigned long long hi, unsigned long long *out) {
    return guarded([&] {
        SQ_THROW_IF(!out || lo > hi || hi > 0x100000000ull, "bad range");
        defaultDevice().makeCurrent();
        *out = expfChecksum(lo, hi);
    });
}

}  // extern "C"
