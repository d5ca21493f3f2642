/*
 * Formulas objects of the rocm device: the DenseGraphFormulas /
 * BipartiteGraphFormulas service of sqaodc (the method table of
 * sqaodc/pyglue/formulas.inc:586-606; CPU forms in
 * sqaodc/cpu/SharedFormulas.cpp, CUDA forms in sqaodc/cuda/DeviceFormulas.cpp).
 *
 * Hamiltonians are formed on the host in double and rounded to real once, as
 * the solvers' setQUBO does; every energy is the batched bilinear form of
 * energy.hip on the device (FP64 accumulation, one rounding to real):
 *   dense QUBO       E = x^T W x                               (SharedFormulas.cpp:10-32)
 *   dense Ising      E = -c - h.q - q^T J q                     (:57-83)
 *   bipartite QUBO   E = b0.x0 + b1.x1 + x1^T W x0              (:86-145)
 *   bipartite Ising  E = -(h0.q0 + h1.q1 + q1^T J q0) - c       (:170-202)
 */
#include <cstring>

#include "solver.h"

namespace sqaod_rocm {

template <class real>
void Formulas<real>::assignDevice(Device &dev) {
    SQ_THROW_IF(dev_ != nullptr && dev_ != &dev, "Device assigned more than once.");
    dev_ = &dev;
}

template <class real>
Device &Formulas<real>::device() {
    if (!dev_)
        dev_ = &defaultDevice();
    dev_->makeCurrent();
    return *dev_;
}

template <class real>
void Formulas<real>::run(QuadFormArgs<real> a, const real *A, size_t nA, const real *lin0, size_t n0,
                         const real *lin1, size_t n1, const int8_t *U, size_t nU, const int8_t *V, size_t nV,
                         real *E, int nOut) {
    Device &dev = device();
    hipStream_t st = dev.stream();
    DevBuf<real> dA, d0, d1, dE;
    DevBuf<int8_t> dU, dV;
    dA.allocate(nA);
    dE.allocate((size_t)nOut);
    dU.allocate(nU);
    HIP_CHECK(hipMemcpyAsync(dA.ptr, A, nA * sizeof(real), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(dU.ptr, U, nU, hipMemcpyHostToDevice, st));
    a.A = dA.ptr;
    a.U = dU.ptr;
    a.V = dU.ptr;
    if (V != U) {
        dV.allocate(nV);
        HIP_CHECK(hipMemcpyAsync(dV.ptr, V, nV, hipMemcpyHostToDevice, st));
        a.V = dV.ptr;
    }
    a.lin0 = a.lin1 = nullptr;
    if (lin0) {
        d0.allocate(n0);
        HIP_CHECK(hipMemcpyAsync(d0.ptr, lin0, n0 * sizeof(real), hipMemcpyHostToDevice, st));
        a.lin0 = d0.ptr;
    }
    if (lin1) {
        d1.allocate(n1);
        HIP_CHECK(hipMemcpyAsync(d1.ptr, lin1, n1 * sizeof(real), hipMemcpyHostToDevice, st));
        a.lin1 = d1.ptr;
    }
    a.E = dE.ptr;
    quadForm<real>(a, nOut, st);
    HIP_CHECK(hipMemcpyAsync(E, dE.ptr, (size_t)nOut * sizeof(real), hipMemcpyDeviceToHost, st));
    dev.synchronize();
}

template <class real>
void Formulas<real>::denseCalculateE(real *E, const real *W, int N, const int8_t *x, int nBatch) {
    /* DGFuncs::calculate_E(E, W, x), SharedFormulas.cpp:10-32 */
    SQ_THROW_IF(N <= 0 || nBatch <= 0, "bad shape N=%d, batch=%d", N, nBatch);
    QuadFormArgs<real> a{};
    a.cst = 0.;
    a.sgn = 1.;
    a.nu = a.nv = N;
    a.nV = nBatch;
    run(a, W, (size_t)N * N, nullptr, 0, nullptr, 0, x, (size_t)N * nBatch, x, 0, E, nBatch);
}

template <class real>
void Formulas<real>::denseHamiltonian(real *h, real *J, real *c, const real *W, int N) {
    /* DGFuncs::calculateHamiltonian, SharedFormulas.cpp:36-55: h = -1/2 colsum W,
     * J = -1/4 W, c = sum J + trace J, then diag J = 0; sums in double */
    SQ_THROW_IF(N <= 0, "W must be a non-empty square matrix.");
    double jsum = 0., dsum = 0.;
    for (int j = 0; j < N; ++j) {
        double s = 0.;
        for (int i = 0; i < N; ++i)
            s += (double)W[(size_t)i * N + j];
        h[j] = (real)(-0.5 * s);
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const real v = real(-0.25) * W[(size_t)i * N + j];
            jsum += (double)v;
            if (i == j)
                dsum += (double)v;
            J[(size_t)i * N + j] = (i == j) ? real(0.) : v;
        }
    *c = (real)(jsum + dsum);
}

template <class real>
void Formulas<real>::denseCalculateEFromSpin(real *E, const real *h, const real *J, real c, const int8_t *q, int N,
                                             int nBatch) {
    /* DGFuncs::calculate_E(E, h, J, c, q), SharedFormulas.cpp:57-83 */
    SQ_THROW_IF(N <= 0 || nBatch <= 0, "bad shape N=%d, batch=%d", N, nBatch);
    QuadFormArgs<real> a{};
    a.cst = -(double)c;
    a.sgn = -1.;
    a.nu = a.nv = N;
    a.nV = nBatch;
    run(a, J, (size_t)N * N, nullptr, 0, h, (size_t)N, q, (size_t)N * nBatch, q, 0, E, nBatch);
}

template <class real>
void Formulas<real>::bipartiteCalculateE(real *E, const real *b0, const real *b1, const real *W, int N0, int N1,
                                         const int8_t *x0, int nBatch0, const int8_t *x1, int nBatch1, bool twoD) {
    /* BGFuncs::calculate_E (SharedFormulas.cpp:86-145): E = b0.x0 + b1.x1 +
     * x1^T W x0, pairwise (x0[b], x1[b]) or over all pairs (E[b1][b0]) */
    SQ_THROW_IF(N0 <= 0 || N1 <= 0 || nBatch0 <= 0 || nBatch1 <= 0, "bad shape");
    SQ_THROW_IF(!twoD && nBatch0 != nBatch1, "Different batch dims between x0 and x1.");
    QuadFormArgs<real> a{};
    a.cst = 0.;
    a.sgn = 1.;
    a.nu = N1;
    a.nv = N0;
    a.nV = nBatch0;
    a.twoD = twoD ? 1 : 0;
    const int nOut = twoD ? nBatch0 * nBatch1 : nBatch0;
    run(a, W, (size_t)N1 * N0, b0, (size_t)N0, b1, (size_t)N1, x1, (size_t)N1 * nBatch1, x0, (size_t)N0 * nBatch0, E,
        nOut);
}

template <class real>
void Formulas<real>::bipartiteHamiltonian(real *h0, real *h1, real *J, real *c, const real *b0, const real *b1,
                                          const real *W, int N0, int N1) {
    /* BGFuncs::calculateHamiltonian, SharedFormulas.cpp:147-167, sums in double */
    SQ_THROW_IF(N0 <= 0 || N1 <= 0, "bad shape");
    double wsum = 0., b0s = 0., b1s = 0.;
    for (int j = 0; j < N0; ++j) {
        double s = 0.;
        for (int i = 0; i < N1; ++i)
            s += (double)W[(size_t)i * N0 + j];
        h0[j] = (real)(-0.25 * s - 0.5 * (double)b0[j]);
    }
    for (int i = 0; i < N1; ++i) {
        double s = 0.;
        for (int j = 0; j < N0; ++j) {
            s += (double)W[(size_t)i * N0 + j];
            J[(size_t)i * N0 + j] = real(-0.25) * W[(size_t)i * N0 + j];
        }
        wsum += s;
        h1[i] = (real)(-0.25 * s - 0.5 * (double)b1[i]);
    }
    for (int j = 0; j < N0; ++j)
        b0s += (double)b0[j];
    for (int i = 0; i < N1; ++i)
        b1s += (double)b1[i];
    *c = (real)(-0.25 * wsum - 0.5 * (b0s + b1s));
}

template <class real>
void Formulas<real>::bipartiteCalculateEFromSpin(real *E, const real *h0, const real *h1, const real *J, real c,
                                                 int N0, int N1, const int8_t *q0, const int8_t *q1, int nBatch) {
    /* BGFuncs::calculate_E(E, h0, h1, J, c, q0, q1), SharedFormulas.cpp:170-202 */
    SQ_THROW_IF(N0 <= 0 || N1 <= 0 || nBatch <= 0, "bad shape");
    QuadFormArgs<real> a{};
    a.cst = -(double)c;
    a.sgn = -1.;
    a.nu = N1;
    a.nv = N0;
    a.nV = nBatch;
    run(a, J, (size_t)N1 * N0, h0, (size_t)N0, h1, (size_t)N1, q1, (size_t)N1 * nBatch, q0, (size_t)N0 * nBatch, E,
        nBatch);
}

template class Formulas<float>;
template class Formulas<double>;

}  // namespace sqaod_rocm
