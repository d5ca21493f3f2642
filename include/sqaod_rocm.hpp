/*
 * sqaod-rocm C++ interface: the rocm device's solvers behind the virtual
 * interface of sqaod's C++ API, for C++ callers of sqaodc.
 *
 * The reference exposes its solvers as abstract class templates
 * (sqaodc/common/Solver.h:18-252: Solver / Annealer / BFSearcher /
 * DenseGraphSolver / BipartiteGraphSolver / DenseGraphAnnealer /
 * BipartiteGraphAnnealer / DenseGraphBFSearcher / BipartiteGraphBFSearcher
 * <real>; sqaodc/common/Formulas.h: DenseGraphFormulas /
 * BipartiteGraphFormulas<real>) created by per-device factories
 * (sqaodc/sqaodc.h:9-25 for sqaod::cpu, :41-59 for sqaod::cuda) and freed
 * with sqaod::deleteInstance (sqaodc/common/Solver.cpp:8-10); the CUDA
 * device adds a Device (initialize / finalize) and assignDevice
 * (sqaodc/cuda/api.h:16-82).  This header declares the same templates, with
 * the same method names, argument meanings, default arguments and error
 * behaviour (std::runtime_error), in namespace sqaod::rocm, together with
 * the value types those signatures use (MatrixType, VectorType, BitSet,
 * BitSetArray, BitSetPair, PackedBitSet, Preference, ...), so code written
 * against sqaod::cuda switches device by namespace:
 *
 *     sqaod::rocm::Device *dev = sqaod::rocm::newDevice(0);
 *     auto *an = sqaod::rocm::newDenseGraphAnnealer<float>();
 *     an->assignDevice(*dev);
 *     an->setQUBO(W);                       // sqaod::rocm::MatrixType<float>
 *     an->setPreference(sqaod::rocm::pnNumTrotters, 32);
 *     an->seed(1000); an->prepare(); an->randomizeSpin();
 *     for (float G = 5.f; G > 0.01f; G *= 0.97f) an->annealOneStep(G, 50.f);
 *     an->makeSolution();
 *     const sqaod::rocm::VectorType<float> &E = an->get_E();
 *     sqaod::rocm::deleteInstance(an); sqaod::rocm::deleteInstance(dev);
 *
 * Header-only, over the C ABI of sqaod_rocm.h (link with libsqaod_rocm.so).
 * Solvers are single-owner objects, as in the reference: calls on one
 * object must not overlap.
 */
#ifndef SQAOD_ROCM_HPP
#define SQAOD_ROCM_HPP

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sqaod_rocm.h"

namespace sqaod {
namespace rocm {

/* ---------------- value types (sqaodc/common/types.h, Matrix.h, Array.h) ---------------- */

typedef int SizeType;
typedef int IdxType;
/* x as bits, element i <-> bit N-1-i (sqaodc/common/Common.cpp:79-93) */
typedef unsigned long long PackedBitSet;

struct NullBase {
    virtual ~NullBase() {}
};

/* sqaodc/common/Preference.h:8-17 (same values) */
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

/* sqaodc/common/Preference.h:26-36 (same values) */
enum PreferenceName {
    pnUnknown = 0,
    pnAlgorithm = 1,
    pnNumTrotters = 2,
    pnTileSize = 3,
    pnTileSize0 = 4,
    pnTileSize1 = 5,
    pnPrecision = 6,
    pnDevice = 7,
    pnExperiment = 100,
};

/* sqaodc/common/Solver.h:9-14 */
enum OptimizeMethod {
    optNone = -1,
    optMinimize = 0,
    optMaximize = 1,
};

inline const char *algorithmToString(Algorithm a) {
    switch (a) {
    case algoDefault: return "default";
    case algoNaive: return "naive";
    case algoColoring: return "coloring";
    case algoBruteForceSearch: return "brute_force_search";
    case algoSADefault: return "sa_default";
    case algoSANaive: return "sa_naive";
    case algoSAColoring: return "sa_coloring";
    default: return "unknown";
    }
}

inline Algorithm algorithmFromString(const std::string &s) {
    for (int a = algoDefault; a <= algoSAColoring; ++a)
        if (s == algorithmToString((Algorithm)a))
            return (Algorithm)a;
    return algoUnknown;
}

inline const char *preferenceNameToString(PreferenceName pn) {
    switch (pn) {
    case pnAlgorithm: return "algorithm";
    case pnNumTrotters: return "n_trotters";
    case pnTileSize: return "tile_size";
    case pnTileSize0: return "tile_size_0";
    case pnTileSize1: return "tile_size_1";
    case pnPrecision: return "precision";
    case pnDevice: return "device";
    case pnExperiment: return "experiment";
    default: return "unknown";
    }
}

inline PreferenceName preferenceNameFromString(const std::string &s) {
    static const PreferenceName all[] = {pnAlgorithm, pnNumTrotters, pnTileSize, pnTileSize0,
                                         pnTileSize1, pnPrecision,   pnDevice,   pnExperiment};
    for (PreferenceName pn : all)
        if (s == preferenceNameToString(pn))
            return pn;
    return pnUnknown;
}

/* A named preference: a size (n_trotters, tile sizes), an algorithm, or a
 * string (precision, device), as sqaodc/common/Preference.h:42-61 holds it. */
struct Preference {
    Preference() : name(pnUnknown), size(0), algo(algoUnknown) {}
    Preference(PreferenceName n, SizeType v) : name(n), size(v), algo(algoUnknown) {}
    Preference(PreferenceName n, Algorithm a) : name(n), size(0), algo(a) {}
    Preference(PreferenceName n, const char *s) : name(n), size(0), algo(algoUnknown), str(s ? s : "") {}
    /* a rocm-device knob by name ("rng", "sweep", "field_precision", ...,
     * the list in sqaod_rocm.h); name stays pnUnknown */
    Preference(const char *knob, const char *value)
        : name(pnUnknown), size(0), algo(algoUnknown), str(value ? value : ""), key(knob ? knob : "") {}
    PreferenceName name;
    SizeType size;
    Algorithm algo;
    std::string str;
    std::string key;
};
typedef std::vector<Preference> Preferences;

/* A row-major matrix, owning its storage or mapping the caller's
 * (stride = row pitch in elements), sqaodc/common/Matrix.h:13-170. */
template <class V>
struct MatrixType {
    typedef V ValueType;
    MatrixType() : rows(0), cols(0), stride(0), data(nullptr) {}
    MatrixType(SizeType r, SizeType c) : rows(r), cols(c), stride(c), own_((size_t)r * c), data(own_.data()) {}
    MatrixType(V *mapped, SizeType r, SizeType c, SizeType s = -1)
        : rows(r), cols(c), stride(s < 0 ? c : s), data(mapped) {}
    MatrixType(const MatrixType &o) : rows(0), cols(0), stride(0), data(nullptr) { *this = o; }
    MatrixType &operator=(const MatrixType &o) {
        if (this != &o) {
            resize(o.rows, o.cols);
            for (SizeType r = 0; r < rows; ++r)
                std::memcpy(rowPtr(r), o.rowPtr(r), (size_t)cols * sizeof(V));
        }
        return *this;
    }
    void resize(SizeType r, SizeType c) {
        if (r == rows && c == cols && data)
            return;
        own_.assign((size_t)r * c, V());
        rows = r;
        cols = c;
        stride = c;
        data = own_.data();
    }
    V &operator()(IdxType r, IdxType c) { return data[(size_t)r * stride + c]; }
    const V &operator()(IdxType r, IdxType c) const { return data[(size_t)r * stride + c]; }
    V *rowPtr(IdxType r) { return data + (size_t)r * stride; }
    const V *rowPtr(IdxType r) const { return data + (size_t)r * stride; }
    /* a dense row-major copy (stride == cols), for the C ABI */
    std::vector<V> packed() const {
        std::vector<V> out((size_t)rows * cols);
        for (SizeType r = 0; r < rows; ++r)
            std::memcpy(&out[(size_t)r * cols], rowPtr(r), (size_t)cols * sizeof(V));
        return out;
    }

    SizeType rows, cols, stride;

private:
    std::vector<V> own_;

public:
    V *data;
};

/* A vector, owning or mapping, sqaodc/common/Matrix.h:215-338. */
template <class V>
struct VectorType {
    typedef V ValueType;
    VectorType() : size(0), data(nullptr) {}
    explicit VectorType(SizeType n) : size(n), own_((size_t)n), data(own_.data()) {}
    VectorType(V *mapped, SizeType n) : size(n), data(mapped) {}
    VectorType(const VectorType &o) : size(0), data(nullptr) { *this = o; }
    VectorType &operator=(const VectorType &o) {
        if (this != &o) {
            resize(o.size);
            if (size)
                std::memcpy(data, o.data, (size_t)size * sizeof(V));
        }
        return *this;
    }
    void resize(SizeType n) {
        if (n == size && data)
            return;
        own_.assign((size_t)n, V());
        size = n;
        data = own_.data();
    }
    V &operator()(IdxType i) { return data[i]; }
    const V &operator()(IdxType i) const { return data[i]; }

    SizeType size;

private:
    std::vector<V> own_;

public:
    V *data;
};

/* sqaodc/common/Array.h:29-150, as a vector with the reference's names */
template <class V>
struct ArrayType : std::vector<V> {
    SizeType size() const { return (SizeType)std::vector<V>::size(); }
    void pushBack(const V &v) { this->push_back(v); }
};

typedef VectorType<char> BitSet;                 /* Matrix.h:369 */
typedef ArrayType<BitSet> BitSetArray;           /* Matrix.h:370 */
typedef std::pair<BitSet, BitSet> BitSetPair;    /* Matrix.h:384-391 */
typedef ArrayType<BitSetPair> BitSetPairArray;   /* Matrix.h:393 */

/* ---------------- the abstract solver API (sqaodc/common/Solver.h:18-252) ---------------- */

template <class real>
struct Solver : NullBase {
    virtual ~Solver() {}
    virtual Algorithm selectAlgorithm(Algorithm algo) = 0;
    virtual Algorithm getAlgorithm() const = 0;
    virtual Preferences getPreferences() const = 0;
    virtual void setPreference(const Preference &pref) = 0;
    template <class V>
    void setPreference(PreferenceName name, const V value) {
        setPreference(Preference(name, value));
    }
    void setPreferences(const Preferences &prefs) {
        for (const Preference &p : prefs)
            setPreference(p);
    }
    virtual const VectorType<real> &get_E() const = 0;
    virtual void prepare() = 0;
    virtual void calculate_E() = 0;
    virtual void makeSolution() = 0;
};

template <class real>
struct BFSearcher : Solver<real> {
    virtual void search() = 0;
};

template <class real>
struct Annealer : Solver<real> {
    using Solver<real>::setPreference;
    virtual void seed(unsigned long long seed) = 0;
    virtual void randomizeSpin() = 0;
    virtual void annealOneStep(real G, real beta) = 0;
    virtual real getSystemE(real G, real beta) const = 0;
};

template <class real>
struct DenseGraphSolver {
    virtual ~DenseGraphSolver() {}
    virtual void getProblemSize(SizeType *N) const = 0;
    virtual void setQUBO(const MatrixType<real> &W, OptimizeMethod om = optMinimize) = 0;
    virtual const BitSetArray &get_x() const = 0;
};

template <class real>
struct BipartiteGraphSolver {
    virtual ~BipartiteGraphSolver() {}
    virtual void getProblemSize(SizeType *N0, SizeType *N1) const = 0;
    virtual void setQUBO(const VectorType<real> &b0, const VectorType<real> &b1, const MatrixType<real> &W,
                         OptimizeMethod om = optMinimize) = 0;
    virtual const BitSetPairArray &get_x() const = 0;
};

template <class real>
struct DenseGraphBFSearcher : BFSearcher<real>, DenseGraphSolver<real> {
    /* scans the next span of x; true once all 2^N are scanned (*curXEnd: the span's end) */
    virtual bool searchRange(PackedBitSet *curXEnd) = 0;
};

template <class real>
struct DenseGraphAnnealer : Annealer<real>, DenseGraphSolver<real> {
    virtual void setHamiltonian(const VectorType<real> &h, const MatrixType<real> &J, real c = real(0.)) = 0;
    virtual void getHamiltonian(VectorType<real> *h, MatrixType<real> *J, real *c) const = 0;
    virtual void set_q(const BitSet &q) = 0;
    virtual void set_qset(const BitSetArray &q) = 0;
    virtual const BitSetArray &get_q() const = 0;
};

template <class real>
struct BipartiteGraphBFSearcher : BFSearcher<real>, BipartiteGraphSolver<real> {
    virtual bool searchRange(PackedBitSet *curX0End, PackedBitSet *curX1End) = 0;
};

template <class real>
struct BipartiteGraphAnnealer : Annealer<real>, BipartiteGraphSolver<real> {
    virtual void setHamiltonian(const VectorType<real> &h0, const VectorType<real> &h1, const MatrixType<real> &J,
                                real c = real(0.)) = 0;
    virtual void getHamiltonian(VectorType<real> *h0, VectorType<real> *h1, MatrixType<real> *J,
                                real *c) const = 0;
    virtual void set_q(const BitSetPair &qPair) = 0;
    virtual void set_qset(const BitSetPairArray &qPairs) = 0;
    virtual const BitSetPairArray &get_q() const = 0;
};

/* sqaodc/common/Formulas.h (x and q as real vectors / matrices, one row per x) */
template <class real>
struct DenseGraphFormulas : NullBase {
    typedef MatrixType<real> Matrix;
    typedef VectorType<real> Vector;
    virtual void calculate_E(real *E, const Matrix &W, const Vector &x) = 0;
    virtual void calculate_E(Vector *E, const Matrix &W, const Matrix &x) = 0;
    virtual void calculateHamiltonian(Vector *h, Matrix *J, real *c, const Matrix &W) = 0;
    virtual void calculate_E(real *E, const Vector &h, const Matrix &J, real c, const Vector &q) = 0;
    virtual void calculate_E(Vector *E, const Vector &h, const Matrix &J, real c, const Matrix &q) = 0;
};

template <class real>
struct BipartiteGraphFormulas : NullBase {
    typedef MatrixType<real> Matrix;
    typedef VectorType<real> Vector;
    virtual void calculate_E(real *E, const Vector &b0, const Vector &b1, const Matrix &W, const Vector &x0,
                             const Vector &x1) = 0;
    virtual void calculate_E(Vector *E, const Vector &b0, const Vector &b1, const Matrix &W, const Matrix &x0,
                             const Matrix &x1) = 0;
    virtual void calculate_E_2d(Matrix *E, const Vector &b0, const Vector &b1, const Matrix &W, const Matrix &x0,
                                const Matrix &x1) = 0;
    virtual void calculateHamiltonian(Vector *h0, Vector *h1, Matrix *J, real *c, const Vector &b0,
                                      const Vector &b1, const Matrix &W) = 0;
    virtual void calculate_E(real *E, const Vector &h0, const Vector &h1, const Matrix &J, real c,
                             const Vector &q0, const Vector &q1) = 0;
    virtual void calculate_E(Vector *E, const Vector &h0, const Vector &h1, const Matrix &J, real c,
                             const Matrix &q0, const Matrix &q1) = 0;
};

/* ---------------- device (sqaodc/cuda/api.h:16-82) ---------------- */

namespace detail {

inline void check(int status) {
    if (status != 0)
        throw std::runtime_error(sqaod_rocm_last_error());
}

template <class real>
constexpr int dtypeOf() {
    static_assert(sizeof(real) == 4 || sizeof(real) == 8, "real is float or double");
    return sizeof(real) == 4 ? 0 : 1;
}

/* "name=value;..." -> pairs */
inline std::vector<std::pair<std::string, std::string>> splitPrefs(const std::string &s) {
    std::vector<std::pair<std::string, std::string>> out;
    size_t b = 0;
    while (b < s.size()) {
        size_t e = s.find(';', b);
        if (e == std::string::npos)
            e = s.size();
        const std::string item = s.substr(b, e - b);
        const size_t eq = item.find('=');
        if (eq != std::string::npos)
            out.emplace_back(item.substr(0, eq), item.substr(eq + 1));
        b = e + 1;
    }
    return out;
}

template <class F>
std::string readPrefs(F get) {
    size_t need = 0;
    check(get(nullptr, 0, &need));
    std::string buf(need, '\0');
    check(get(&buf[0], buf.size(), &need));
    buf.resize(std::strlen(buf.c_str()));
    return buf;
}

inline Preferences toPreferences(const std::string &s) {
    Preferences out;
    for (auto &kv : splitPrefs(s)) {
        const PreferenceName pn = preferenceNameFromString(kv.first);
        if (pn == pnAlgorithm)
            out.emplace_back(pn, algorithmFromString(kv.second));
        else if (pn == pnNumTrotters || pn == pnTileSize || pn == pnTileSize0 || pn == pnTileSize1)
            out.emplace_back(pn, (SizeType)std::stoi(kv.second));
        else if (pn == pnPrecision || pn == pnDevice)
            out.emplace_back(pn, kv.second.c_str());
    }
    return out;
}

/* the C ABI's (name, value) text of a Preference */
inline std::pair<std::string, std::string> prefText(const Preference &p) {
    const std::string name = preferenceNameToString(p.name);
    if (p.name == pnAlgorithm)
        return {name, algorithmToString(p.algo)};
    if (p.name == pnPrecision || p.name == pnDevice)
        return {name, p.str};
    if (p.name == pnUnknown) {
        if (p.key.empty())
            throw std::runtime_error("unknown preference");
        return {p.key, p.str};
    }
    return {name, std::to_string(p.size)};
}

inline std::vector<int8_t> bitsOf(const BitSet &b) {
    std::vector<int8_t> out((size_t)b.size);
    for (SizeType i = 0; i < b.size; ++i)
        out[i] = (int8_t)b.data[i];
    return out;
}

inline BitSetArray rowsOf(const std::vector<int8_t> &v, int n, int N) {
    BitSetArray out;
    for (int r = 0; r < n; ++r) {
        BitSet b(N);
        for (int i = 0; i < N; ++i)
            b.data[i] = (char)v[(size_t)r * N + i];
        out.pushBack(b);
    }
    return out;
}

template <class real>
std::vector<int8_t> int8Rows(const MatrixType<real> &x) {
    std::vector<int8_t> out((size_t)x.rows * x.cols);
    for (SizeType r = 0; r < x.rows; ++r)
        for (SizeType c = 0; c < x.cols; ++c)
            out[(size_t)r * x.cols + c] = (int8_t)x(r, c);
    return out;
}

template <class real>
std::vector<int8_t> int8Vec(const VectorType<real> &x) {
    std::vector<int8_t> out((size_t)x.size);
    for (SizeType i = 0; i < x.size; ++i)
        out[i] = (int8_t)x.data[i];
    return out;
}

}  // namespace detail

class Device : public NullBase {
public:
    Device() : h_(nullptr) {}
    ~Device() override { finalize(); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    /* attach to GPU devNo (a HIP stream of its own; cuda/Device.cpp initialize) */
    virtual void initialize(int devNo = 0) {
        finalize();
        detail::check(sqaod_rocm_device_new(devNo, &h_));
    }
    virtual void finalize() {
        if (h_)
            (void)sqaod_rocm_device_delete(h_);
        h_ = nullptr;
    }
    void synchronize() { detail::check(sqaod_rocm_device_synchronize(h_)); }
    sqaod_rocm_device *handle() const {
        if (!h_)
            throw std::runtime_error("Device not initialized");
        return h_;
    }

private:
    sqaod_rocm_device *h_;
};

/* sqaodc/cuda/api.h DeviceAssigner */
struct DeviceAssigner {
    virtual ~DeviceAssigner() {}
    virtual void assignDevice(Device &device) = 0;
};

/* ---------------- implementations over the C ABI ---------------- */

namespace detail {

template <class real>
class DenseGraphAnnealerImpl : public rocm::DenseGraphAnnealer<real>, public DeviceAssigner {
public:
    DenseGraphAnnealerImpl() { check(sqaod_rocm_dg_annealer_new(dtypeOf<real>(), &h_)); }
    ~DenseGraphAnnealerImpl() override { (void)sqaod_rocm_dg_annealer_delete(h_); }
    void assignDevice(Device &d) override { check(sqaod_rocm_dg_annealer_assign_device(h_, d.handle())); }
    void seed(unsigned long long s) override { check(sqaod_rocm_dg_annealer_seed(h_, s)); }
    Algorithm selectAlgorithm(Algorithm a) override {
        check(sqaod_rocm_dg_annealer_set_preference(h_, "algorithm", algorithmToString(a)));
        return getAlgorithm();
    }
    Algorithm getAlgorithm() const override {
        for (auto &kv : splitPrefs(prefText()))
            if (kv.first == "algorithm")
                return algorithmFromString(kv.second);
        return algoUnknown;
    }
    Preferences getPreferences() const override { return toPreferences(prefText()); }
    void setPreference(const Preference &p) override {
        auto t = detail::prefText(p);
        check(sqaod_rocm_dg_annealer_set_preference(h_, t.first.c_str(), t.second.c_str()));
    }
    using rocm::DenseGraphAnnealer<real>::setPreference;
    void getProblemSize(SizeType *N) const override { check(sqaod_rocm_dg_annealer_get_problem_size(h_, N)); }
    void setQUBO(const MatrixType<real> &W, OptimizeMethod om = optMinimize) override {
        if (W.rows != W.cols)
            throw std::runtime_error("W is not a square matrix.");
        const std::vector<real> w = W.packed();
        check(sqaod_rocm_dg_annealer_set_qubo(h_, w.data(), W.rows, om == optMaximize ? 1 : 0));
    }
    void setHamiltonian(const VectorType<real> &h, const MatrixType<real> &J, real c = real(0.)) override {
        if (J.rows != J.cols || h.size != J.rows)
            throw std::runtime_error("Shape mismatch: h[N], J[N x N].");
        const std::vector<real> j = J.packed();
        check(sqaod_rocm_dg_annealer_set_hamiltonian(h_, h.data, j.data(), &c, J.rows));
    }
    void getHamiltonian(VectorType<real> *h, MatrixType<real> *J, real *c) const override {
        SizeType N = 0;
        getProblemSize(&N);
        h->resize(N);
        J->resize(N, N);
        check(sqaod_rocm_dg_annealer_get_hamiltonian(h_, h->data, J->data, c));
    }
    void set_q(const BitSet &q) override {
        const std::vector<int8_t> b = bitsOf(q);
        check(sqaod_rocm_dg_annealer_set_q(h_, b.data(), q.size));
    }
    void set_qset(const BitSetArray &q) override {
        if (q.empty())
            throw std::runtime_error("empty q set");
        const int N = q[0].size;
        std::vector<int8_t> b;
        for (const BitSet &r : q) {
            if (r.size != N)
                throw std::runtime_error("Shape mismatch in q set.");
            const std::vector<int8_t> v = bitsOf(r);
            b.insert(b.end(), v.begin(), v.end());
        }
        check(sqaod_rocm_dg_annealer_set_qset(h_, b.data(), q.size(), N));
    }
    const BitSetArray &get_q() const override {
        int m = 0, N = 0;
        shape(&m, &N);
        std::vector<int8_t> v((size_t)m * N);
        check(sqaod_rocm_dg_annealer_get_q(h_, v.data(), m, N));
        q_ = rowsOf(v, m, N);
        return q_;
    }
    const BitSetArray &get_x() const override {
        int m = 0, N = 0;
        shape(&m, &N);
        std::vector<int8_t> v((size_t)m * N);
        check(sqaod_rocm_dg_annealer_get_x(h_, v.data(), m, N));
        x_ = rowsOf(v, m, N);
        return x_;
    }
    const VectorType<real> &get_E() const override {
        int m = 0, N = 0;
        shape(&m, &N);
        E_.resize(m);
        check(sqaod_rocm_dg_annealer_get_E(h_, E_.data, m));
        return E_;
    }
    void randomizeSpin() override { check(sqaod_rocm_dg_annealer_randomize_spin(h_)); }
    void prepare() override { check(sqaod_rocm_dg_annealer_prepare(h_)); }
    void calculate_E() override { check(sqaod_rocm_dg_annealer_calculate_E(h_)); }
    void makeSolution() override { check(sqaod_rocm_dg_annealer_make_solution(h_)); }
    real getSystemE(real G, real beta) const override {
        double E = 0.;
        check(sqaod_rocm_dg_annealer_get_system_E(h_, (double)G, (double)beta, &E));
        return (real)E;
    }
    void annealOneStep(real G, real beta) override {
        check(sqaod_rocm_dg_annealer_anneal_one_step(h_, (double)G, (double)beta));
    }
    sqaod_rocm_dg_annealer *handle() const { return h_; }

private:
    std::string prefText() const {
        return readPrefs([&](char *b, size_t n, size_t *need) {
            return sqaod_rocm_dg_annealer_get_preferences(h_, b, n, need);
        });
    }
    void shape(int *m, int *N) const {
        check(sqaod_rocm_dg_annealer_get_n_trotters(h_, m));
        check(sqaod_rocm_dg_annealer_get_problem_size(h_, N));
    }
    sqaod_rocm_dg_annealer *h_ = nullptr;
    mutable BitSetArray q_, x_;
    mutable VectorType<real> E_;
};

template <class real>
class BipartiteGraphAnnealerImpl : public rocm::BipartiteGraphAnnealer<real>, public DeviceAssigner {
public:
    BipartiteGraphAnnealerImpl() { check(sqaod_rocm_bg_annealer_new(dtypeOf<real>(), &h_)); }
    ~BipartiteGraphAnnealerImpl() override { (void)sqaod_rocm_bg_annealer_delete(h_); }
    void assignDevice(Device &d) override { check(sqaod_rocm_bg_annealer_assign_device(h_, d.handle())); }
    void seed(unsigned long long s) override { check(sqaod_rocm_bg_annealer_seed(h_, s)); }
    Algorithm selectAlgorithm(Algorithm a) override {
        check(sqaod_rocm_bg_annealer_set_preference(h_, "algorithm", algorithmToString(a)));
        return getAlgorithm();
    }
    Algorithm getAlgorithm() const override {
        for (auto &kv : splitPrefs(prefText()))
            if (kv.first == "algorithm")
                return algorithmFromString(kv.second);
        return algoUnknown;
    }
    Preferences getPreferences() const override { return toPreferences(prefText()); }
    void setPreference(const Preference &p) override {
        auto t = detail::prefText(p);
        check(sqaod_rocm_bg_annealer_set_preference(h_, t.first.c_str(), t.second.c_str()));
    }
    using rocm::BipartiteGraphAnnealer<real>::setPreference;
    void getProblemSize(SizeType *N0, SizeType *N1) const override {
        check(sqaod_rocm_bg_annealer_get_problem_size(h_, N0, N1));
    }
    void setQUBO(const VectorType<real> &b0, const VectorType<real> &b1, const MatrixType<real> &W,
                 OptimizeMethod om = optMinimize) override {
        if (W.rows != b1.size || W.cols != b0.size)
            throw std::runtime_error("Shape mismatch: W must be N1 x N0 for b0[N0], b1[N1].");
        const std::vector<real> w = W.packed();
        check(sqaod_rocm_bg_annealer_set_qubo(h_, b0.data, b0.size, b1.data, b1.size, w.data(),
                                              om == optMaximize ? 1 : 0));
    }
    void setHamiltonian(const VectorType<real> &h0, const VectorType<real> &h1, const MatrixType<real> &J,
                        real c = real(0.)) override {
        if (J.rows != h1.size || J.cols != h0.size)
            throw std::runtime_error("Shape mismatch: J must be N1 x N0 for h0[N0], h1[N1].");
        const std::vector<real> j = J.packed();
        check(sqaod_rocm_bg_annealer_set_hamiltonian(h_, h0.data, h0.size, h1.data, h1.size, j.data(), &c));
    }
    void getHamiltonian(VectorType<real> *h0, VectorType<real> *h1, MatrixType<real> *J,
                        real *c) const override {
        SizeType N0 = 0, N1 = 0;
        getProblemSize(&N0, &N1);
        h0->resize(N0);
        h1->resize(N1);
        J->resize(N1, N0);
        check(sqaod_rocm_bg_annealer_get_hamiltonian(h_, h0->data, h1->data, J->data, c));
    }
    void set_q(const BitSetPair &qp) override {
        const std::vector<int8_t> a = bitsOf(qp.first), b = bitsOf(qp.second);
        check(sqaod_rocm_bg_annealer_set_q(h_, a.data(), qp.first.size, b.data(), qp.second.size));
    }
    void set_qset(const BitSetPairArray &qs) override {
        if (qs.empty())
            throw std::runtime_error("empty q set");
        const int N0 = qs[0].first.size, N1 = qs[0].second.size;
        std::vector<int8_t> a, b;
        for (const BitSetPair &p : qs) {
            if (p.first.size != N0 || p.second.size != N1)
                throw std::runtime_error("Shape mismatch in q set.");
            const std::vector<int8_t> u = bitsOf(p.first), v = bitsOf(p.second);
            a.insert(a.end(), u.begin(), u.end());
            b.insert(b.end(), v.begin(), v.end());
        }
        check(sqaod_rocm_bg_annealer_set_qset(h_, a.data(), b.data(), qs.size(), N0, N1));
    }
    const BitSetPairArray &get_q() const override {
        q_ = pairs(sqaod_rocm_bg_annealer_get_q);
        return q_;
    }
    const BitSetPairArray &get_x() const override {
        x_ = pairs(sqaod_rocm_bg_annealer_get_x);
        return x_;
    }
    const VectorType<real> &get_E() const override {
        int m = 0;
        check(sqaod_rocm_bg_annealer_get_n_trotters(h_, &m));
        E_.resize(m);
        check(sqaod_rocm_bg_annealer_get_E(h_, E_.data, m));
        return E_;
    }
    void randomizeSpin() override { check(sqaod_rocm_bg_annealer_randomize_spin(h_)); }
    void prepare() override { check(sqaod_rocm_bg_annealer_prepare(h_)); }
    void calculate_E() override { check(sqaod_rocm_bg_annealer_calculate_E(h_)); }
    void makeSolution() override { check(sqaod_rocm_bg_annealer_make_solution(h_)); }
    real getSystemE(real G, real beta) const override {
        double E = 0.;
        check(sqaod_rocm_bg_annealer_get_system_E(h_, (double)G, (double)beta, &E));
        return (real)E;
    }
    void annealOneStep(real G, real beta) override {
        check(sqaod_rocm_bg_annealer_anneal_one_step(h_, (double)G, (double)beta));
    }

private:
    template <class F>
    BitSetPairArray pairs(F get) const {
        int m = 0;
        SizeType N0 = 0, N1 = 0;
        check(sqaod_rocm_bg_annealer_get_n_trotters(h_, &m));
        getProblemSize(&N0, &N1);
        std::vector<int8_t> a((size_t)m * N0), b((size_t)m * N1);
        check(get(h_, a.data(), b.data(), m, N0, N1));
        BitSetArray r0 = rowsOf(a, m, N0), r1 = rowsOf(b, m, N1);
        BitSetPairArray out;
        for (int y = 0; y < m; ++y)
            out.pushBack(BitSetPair(r0[y], r1[y]));
        return out;
    }
    std::string prefText() const {
        return readPrefs([&](char *b, size_t n, size_t *need) {
            return sqaod_rocm_bg_annealer_get_preferences(h_, b, n, need);
        });
    }
    sqaod_rocm_bg_annealer *h_ = nullptr;
    mutable BitSetPairArray q_, x_;
    mutable VectorType<real> E_;
};

template <class real>
class DenseGraphBFSearcherImpl : public rocm::DenseGraphBFSearcher<real>, public DeviceAssigner {
public:
    DenseGraphBFSearcherImpl() { check(sqaod_rocm_dg_bf_searcher_new(dtypeOf<real>(), &h_)); }
    ~DenseGraphBFSearcherImpl() override { (void)sqaod_rocm_dg_bf_searcher_delete(h_); }
    void assignDevice(Device &d) override { check(sqaod_rocm_dg_bf_searcher_assign_device(h_, d.handle())); }
    /* BFSearcher::selectAlgorithm (Solver.cpp:179-190): always brute force */
    Algorithm selectAlgorithm(Algorithm) override { return algoBruteForceSearch; }
    Algorithm getAlgorithm() const override { return algoBruteForceSearch; }
    Preferences getPreferences() const override {
        return toPreferences(readPrefs([&](char *b, size_t n, size_t *need) {
            return sqaod_rocm_dg_bf_searcher_get_preferences(h_, b, n, need);
        }));
    }
    void setPreference(const Preference &p) override {
        auto t = detail::prefText(p);
        check(sqaod_rocm_dg_bf_searcher_set_preference(h_, t.first.c_str(), t.second.c_str()));
    }
    using rocm::DenseGraphBFSearcher<real>::setPreference;
    void getProblemSize(SizeType *N) const override { check(sqaod_rocm_dg_bf_searcher_get_problem_size(h_, N)); }
    void setQUBO(const MatrixType<real> &W, OptimizeMethod om = optMinimize) override {
        if (W.rows != W.cols)
            throw std::runtime_error("W is not a square matrix.");
        const std::vector<real> w = W.packed();
        check(sqaod_rocm_dg_bf_searcher_set_qubo(h_, w.data(), W.rows, om == optMaximize ? 1 : 0));
    }
    const BitSetArray &get_x() const override {
        int n = 0;
        SizeType N = 0;
        check(sqaod_rocm_dg_bf_searcher_get_n_solutions(h_, &n));
        getProblemSize(&N);
        std::vector<int8_t> v((size_t)n * N);
        check(sqaod_rocm_dg_bf_searcher_get_x(h_, v.data(), n, N));
        x_ = rowsOf(v, n, N);
        return x_;
    }
    const VectorType<real> &get_E() const override {
        int n = 0;
        check(sqaod_rocm_dg_bf_searcher_get_n_solutions(h_, &n));
        n = n < 1 ? 1 : n;
        E_.resize(n);
        check(sqaod_rocm_dg_bf_searcher_get_E(h_, E_.data, n));
        return E_;
    }
    void prepare() override { check(sqaod_rocm_dg_bf_searcher_prepare(h_)); }
    void calculate_E() override { check(sqaod_rocm_dg_bf_searcher_calculate_E(h_)); }
    void makeSolution() override { check(sqaod_rocm_dg_bf_searcher_make_solution(h_)); }
    bool searchRange(PackedBitSet *curXEnd) override {
        int done = 0;
        unsigned long long end = 0;
        check(sqaod_rocm_dg_bf_searcher_search_range(h_, &done, &end));
        if (curXEnd)
            *curXEnd = end;
        return done != 0;
    }
    void search() override { check(sqaod_rocm_dg_bf_searcher_search(h_)); }

private:
    sqaod_rocm_dg_bf_searcher *h_ = nullptr;
    mutable BitSetArray x_;
    mutable VectorType<real> E_;
};

template <class real>
class BipartiteGraphBFSearcherImpl : public rocm::BipartiteGraphBFSearcher<real>, public DeviceAssigner {
public:
    BipartiteGraphBFSearcherImpl() { check(sqaod_rocm_bg_bf_searcher_new(dtypeOf<real>(), &h_)); }
    ~BipartiteGraphBFSearcherImpl() override { (void)sqaod_rocm_bg_bf_searcher_delete(h_); }
    void assignDevice(Device &d) override { check(sqaod_rocm_bg_bf_searcher_assign_device(h_, d.handle())); }
    Algorithm selectAlgorithm(Algorithm) override { return algoBruteForceSearch; }
    Algorithm getAlgorithm() const override { return algoBruteForceSearch; }
    Preferences getPreferences() const override {
        return toPreferences(readPrefs([&](char *b, size_t n, size_t *need) {
            return sqaod_rocm_bg_bf_searcher_get_preferences(h_, b, n, need);
        }));
    }
    void setPreference(const Preference &p) override {
        auto t = detail::prefText(p);
        check(sqaod_rocm_bg_bf_searcher_set_preference(h_, t.first.c_str(), t.second.c_str()));
    }
    using rocm::BipartiteGraphBFSearcher<real>::setPreference;
    void getProblemSize(SizeType *N0, SizeType *N1) const override {
        check(sqaod_rocm_bg_bf_searcher_get_problem_size(h_, N0, N1));
    }
    void setQUBO(const VectorType<real> &b0, const VectorType<real> &b1, const MatrixType<real> &W,
                 OptimizeMethod om = optMinimize) override {
        if (W.rows != b1.size || W.cols != b0.size)
            throw std::runtime_error("Shape mismatch: W must be N1 x N0 for b0[N0], b1[N1].");
        const std::vector<real> w = W.packed();
        check(sqaod_rocm_bg_bf_searcher_set_qubo(h_, b0.data, b0.size, b1.data, b1.size, w.data(),
                                                 om == optMaximize ? 1 : 0));
    }
    const BitSetPairArray &get_x() const override {
        int n = 0;
        SizeType N0 = 0, N1 = 0;
        check(sqaod_rocm_bg_bf_searcher_get_n_solutions(h_, &n));
        getProblemSize(&N0, &N1);
        std::vector<int8_t> a((size_t)n * N0), b((size_t)n * N1);
        check(sqaod_rocm_bg_bf_searcher_get_x(h_, a.data(), b.data(), n, N0, N1));
        BitSetArray r0 = rowsOf(a, n, N0), r1 = rowsOf(b, n, N1);
        x_.clear();
        for (int i = 0; i < n; ++i)
            x_.pushBack(BitSetPair(r0[i], r1[i]));
        return x_;
    }
    const VectorType<real> &get_E() const override {
        int n = 0;
        check(sqaod_rocm_bg_bf_searcher_get_n_solutions(h_, &n));
        n = n < 1 ? 1 : n;
        E_.resize(n);
        check(sqaod_rocm_bg_bf_searcher_get_E(h_, E_.data, n));
        return E_;
    }
    void prepare() override { check(sqaod_rocm_bg_bf_searcher_prepare(h_)); }
    void calculate_E() override { check(sqaod_rocm_bg_bf_searcher_calculate_E(h_)); }
    void makeSolution() override { check(sqaod_rocm_bg_bf_searcher_make_solution(h_)); }
    bool searchRange(PackedBitSet *curX0End, PackedBitSet *curX1End) override {
        int done = 0;
        unsigned long long x0 = 0, x1 = 0;
        check(sqaod_rocm_bg_bf_searcher_search_range(h_, &done, &x0, &x1));
        if (curX0End)
            *curX0End = x0;
        if (curX1End)
            *curX1End = x1;
        return done != 0;
    }
    void search() override { check(sqaod_rocm_bg_bf_searcher_search(h_)); }

private:
    sqaod_rocm_bg_bf_searcher *h_ = nullptr;
    mutable BitSetPairArray x_;
    mutable VectorType<real> E_;
};

template <class real>
class DenseGraphFormulasImpl : public rocm::DenseGraphFormulas<real>, public DeviceAssigner {
public:
    typedef MatrixType<real> Matrix;
    typedef VectorType<real> Vector;
    DenseGraphFormulasImpl() { check(sqaod_rocm_formulas_new(&h_)); }
    ~DenseGraphFormulasImpl() override { (void)sqaod_rocm_formulas_delete(h_); }
    void assignDevice(Device &d) override { check(sqaod_rocm_formulas_assign_device(h_, d.handle())); }
    void calculate_E(real *E, const Matrix &W, const Vector &x) override {
        const std::vector<real> w = W.packed();
        const std::vector<int8_t> b = int8Vec(x);
        check(sqaod_rocm_dense_graph_batch_calculate_E(h_, dtypeOf<real>(), E, w.data(), W.rows, b.data(), 1));
    }
    void calculate_E(Vector *E, const Matrix &W, const Matrix &x) override {
        const std::vector<real> w = W.packed();
        const std::vector<int8_t> b = int8Rows(x);
        E->resize(x.rows);
        check(sqaod_rocm_dense_graph_batch_calculate_E(h_, dtypeOf<real>(), E->data, w.data(), W.rows, b.data(),
                                                       x.rows));
    }
    void calculateHamiltonian(Vector *h, Matrix *J, real *c, const Matrix &W) override {
        const std::vector<real> w = W.packed();
        h->resize(W.rows);
        J->resize(W.rows, W.rows);
        check(sqaod_rocm_dense_graph_calculate_hamiltonian(h_, dtypeOf<real>(), h->data, J->data, c, w.data(),
                                                           W.rows));
    }
    void calculate_E(real *E, const Vector &hv, const Matrix &J, real c, const Vector &q) override {
        const std::vector<real> j = J.packed();
        const std::vector<int8_t> b = int8Vec(q);
        check(sqaod_rocm_dense_graph_batch_calculate_E_from_spin(h_, dtypeOf<real>(), E, hv.data, j.data(), &c,
                                                                 b.data(), J.rows, 1));
    }
    void calculate_E(Vector *E, const Vector &hv, const Matrix &J, real c, const Matrix &q) override {
        const std::vector<real> j = J.packed();
        const std::vector<int8_t> b = int8Rows(q);
        E->resize(q.rows);
        check(sqaod_rocm_dense_graph_batch_calculate_E_from_spin(h_, dtypeOf<real>(), E->data, hv.data, j.data(),
                                                                 &c, b.data(), J.rows, q.rows));
    }

private:
    sqaod_rocm_formulas *h_ = nullptr;
};

template <class real>
class BipartiteGraphFormulasImpl : public rocm::BipartiteGraphFormulas<real>, public DeviceAssigner {
public:
    typedef MatrixType<real> Matrix;
    typedef VectorType<real> Vector;
    BipartiteGraphFormulasImpl() { check(sqaod_rocm_formulas_new(&h_)); }
    ~BipartiteGraphFormulasImpl() override { (void)sqaod_rocm_formulas_delete(h_); }
    void assignDevice(Device &d) override { check(sqaod_rocm_formulas_assign_device(h_, d.handle())); }
    void calculate_E(real *E, const Vector &b0, const Vector &b1, const Matrix &W, const Vector &x0,
                     const Vector &x1) override {
        const std::vector<real> w = W.packed();
        const std::vector<int8_t> a = int8Vec(x0), b = int8Vec(x1);
        check(sqaod_rocm_bipartite_graph_batch_calculate_E(h_, dtypeOf<real>(), E, b0.data, b0.size, b1.data,
                                                           b1.size, w.data(), a.data(), b.data(), 1));
    }
    void calculate_E(Vector *E, const Vector &b0, const Vector &b1, const Matrix &W, const Matrix &x0,
                     const Matrix &x1) override {
        const std::vector<real> w = W.packed();
        const std::vector<int8_t> a = int8Rows(x0), b = int8Rows(x1);
        if (x0.rows != x1.rows)
            throw std::runtime_error("Shape mismatch: x0 and x1 batches differ in size.");
        E->resize(x0.rows);
        check(sqaod_rocm_bipartite_graph_batch_calculate_E(h_, dtypeOf<real>(), E->data, b0.data, b0.size,
                                                           b1.data, b1.size, w.data(), a.data(), b.data(),
                                                           x0.rows));
    }
    void calculate_E_2d(Matrix *E, const Vector &b0, const Vector &b1, const Matrix &W, const Matrix &x0,
                        const Matrix &x1) override {
        const std::vector<real> w = W.packed();
        const std::vector<int8_t> a = int8Rows(x0), b = int8Rows(x1);
        E->resize(x1.rows, x0.rows);
        check(sqaod_rocm_bipartite_graph_batch_calculate_E_2d(h_, dtypeOf<real>(), E->data, b0.data, b0.size,
                                                              b1.data, b1.size, w.data(), a.data(), x0.rows,
                                                              b.data(), x1.rows));
    }
    void calculateHamiltonian(Vector *h0, Vector *h1, Matrix *J, real *c, const Vector &b0, const Vector &b1,
                              const Matrix &W) override {
        const std::vector<real> w = W.packed();
        h0->resize(b0.size);
        h1->resize(b1.size);
        J->resize(b1.size, b0.size);
        check(sqaod_rocm_bipartite_graph_calculate_hamiltonian(h_, dtypeOf<real>(), h0->data, h1->data, J->data, c,
                                                               b0.data, b0.size, b1.data, b1.size, w.data()));
    }
    void calculate_E(real *E, const Vector &h0, const Vector &h1, const Matrix &J, real c, const Vector &q0,
                     const Vector &q1) override {
        const std::vector<real> j = J.packed();
        const std::vector<int8_t> a = int8Vec(q0), b = int8Vec(q1);
        check(sqaod_rocm_bipartite_graph_batch_calculate_E_from_spin(h_, dtypeOf<real>(), E, h0.data, h0.size,
                                                                     h1.data, h1.size, j.data(), &c, a.data(),
                                                                     b.data(), 1));
    }
    void calculate_E(Vector *E, const Vector &h0, const Vector &h1, const Matrix &J, real c, const Matrix &q0,
                     const Matrix &q1) override {
        const std::vector<real> j = J.packed();
        const std::vector<int8_t> a = int8Rows(q0), b = int8Rows(q1);
        if (q0.rows != q1.rows)
            throw std::runtime_error("Shape mismatch: q0 and q1 batches differ in size.");
        E->resize(q0.rows);
        check(sqaod_rocm_bipartite_graph_batch_calculate_E_from_spin(h_, dtypeOf<real>(), E->data, h0.data,
                                                                     h0.size, h1.data, h1.size, j.data(), &c,
                                                                     a.data(), b.data(), q0.rows));
    }

private:
    sqaod_rocm_formulas *h_ = nullptr;
};

}  // namespace detail

/* ---------------- factories (sqaodc/sqaodc.h:41-59, sqaodc/cuda/api_cuda.cpp:11-69) ---------------- */

template <class real>
detail::DenseGraphAnnealerImpl<real> *newDenseGraphAnnealer() {
    return new detail::DenseGraphAnnealerImpl<real>();
}

template <class real>
detail::BipartiteGraphAnnealerImpl<real> *newBipartiteGraphAnnealer() {
    return new detail::BipartiteGraphAnnealerImpl<real>();
}

template <class real>
detail::DenseGraphBFSearcherImpl<real> *newDenseGraphBFSearcher() {
    return new detail::DenseGraphBFSearcherImpl<real>();
}

template <class real>
detail::BipartiteGraphBFSearcherImpl<real> *newBipartiteGraphBFSearcher() {
    return new detail::BipartiteGraphBFSearcherImpl<real>();
}

template <class real>
detail::DenseGraphFormulasImpl<real> *newDenseGraphFormulas() {
    return new detail::DenseGraphFormulasImpl<real>();
}

template <class real>
detail::BipartiteGraphFormulasImpl<real> *newBipartiteGraphFormulas() {
    return new detail::BipartiteGraphFormulasImpl<real>();
}

/* a Device, initialized when devNo >= 0 (sqaodc/cuda/api_cuda.cpp newDevice) */
inline Device *newDevice(int devNo = -1) {
    Device *d = new Device();
    if (devNo >= 0) {
        try {
            d->initialize(devNo);
        } catch (...) {
            delete d;
            throw;
        }
    }
    return d;
}

/* sqaodc/common/Solver.cpp:8-10 */
inline void deleteInstance(NullBase *obj) { delete obj; }

}  // namespace rocm
}  // namespace sqaod

#endif
