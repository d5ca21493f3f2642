/*
 * Drives the rocm solvers through the C++ interface of include/sqaod_rocm.hpp
 * (the sqaod::rocm factories and the Solver.h-shaped virtual methods) and
 * writes what they return as JSON, for tests/test_cpp_api_gpu.py to check
 * against the oracle:
 *   - a dense anneal, N = 128, m = 32, integer W, the device MT19937 stream,
 *     the 205-step schedule (seed / setQUBO / prepare / randomizeSpin /
 *     annealOneStep / makeSolution / get_E / get_x / get_q / getSystemE);
 *   - a bipartite anneal, 40 x 24, m = 8, integer b0, b1, W, 60 steps;
 *   - a dense brute-force search, N = 12 (searchRange loop + makeSolution);
 *   - the dense formulas: calculate_E of a batch of x.
 * Usage: cpp_api_run OUT.json
 */
#include <cstdio>
#include <string>
#include <vector>

#include "sqaod_rocm.hpp"

namespace sr = sqaod::rocm;

namespace {

/* integer weights in [-8, 8] from a 64-bit LCG (Knuth's MMIX constants) */
struct Lcg {
    unsigned long long s;
    int next() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (int)((s >> 33) % 17) - 8;
    }
};

sr::MatrixType<float> symmetricW(int N, unsigned long long seed) {
    Lcg g{seed};
    sr::MatrixType<float> W(N, N);
    for (int i = 0; i < N; ++i)
        for (int j = i; j < N; ++j)
            W(i, j) = W(j, i) = (float)g.next();
    return W;
}

template <class V>
void putVec(std::string &o, const char *key, const V *v, size_t n) {
    o += "\"";
    o += key;
    o += "\": [";
    for (size_t i = 0; i < n; ++i) {
        char b[64];
        snprintf(b, sizeof b, "%s%.9g", i ? ", " : "", (double)v[i]);
        o += b;
    }
    o += "], ";
}

void putRows(std::string &o, const char *key, const sr::BitSetArray &rows) {
    std::vector<int> flat;
    for (const sr::BitSet &r : rows)
        for (int i = 0; i < r.size; ++i)
            flat.push_back(r.data[i]);
    putVec(o, key, flat.data(), flat.size());
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s OUT.json\n", argv[0]);
        return 2;
    }
    std::string o = "{";
    try {
        sr::Device *dev = sr::newDevice(0);

        /* dense annealer */
        const int N = 128, m = 32;
        sr::MatrixType<float> W = symmetricW(N, 1133557);
        auto *an = sr::newDenseGraphAnnealer<float>();
        an->assignDevice(*dev);
        an->setQUBO(W);
        an->setPreference(sr::pnNumTrotters, m);
        an->setPreference(sr::Preference("rng", "mt19937"));
        an->selectAlgorithm(sr::algoColoring);
        an->seed(1000);
        an->prepare();
        an->randomizeSpin();
        std::vector<float> G;
        for (double g = 5.0; g > 0.01; g *= 0.97)
            G.push_back((float)g);
        for (float g : G)
            an->annealOneStep(g, 50.f);
        an->makeSolution();
        const sr::VectorType<float> &E = an->get_E();
        sr::SizeType n = 0;
        an->getProblemSize(&n);
        const float sysE = an->getSystemE(G.back(), 50.f);
        putVec(o, "dense_W", W.data, (size_t)N * N);
        putVec(o, "dense_G", G.data(), G.size());
        putVec(o, "dense_E", E.data, (size_t)E.size);
        putRows(o, "dense_x", an->get_x());
        putRows(o, "dense_q", an->get_q());
        putVec(o, "dense_systemE", &sysE, 1);
        const int shape[3] = {n, m, (int)an->getAlgorithm()};
        putVec(o, "dense_shape", shape, 3);
        int nt = -1;
        for (const sr::Preference &p : an->getPreferences())
            if (p.name == sr::pnNumTrotters)
                nt = p.size;
        putVec(o, "dense_pref_n_trotters", &nt, 1);
        sr::deleteInstance(an);

        /* bipartite annealer */
        const int N0 = 40, N1 = 24, mb = 8;
        Lcg g{77};
        sr::VectorType<float> b0(N0), b1(N1);
        sr::MatrixType<float> Wb(N1, N0);
        for (int i = 0; i < N0; ++i)
            b0(i) = (float)g.next();
        for (int i = 0; i < N1; ++i)
            b1(i) = (float)g.next();
        for (int i = 0; i < N1; ++i)
            for (int j = 0; j < N0; ++j)
                Wb(i, j) = (float)g.next();
        auto *bg = sr::newBipartiteGraphAnnealer<float>();
        bg->assignDevice(*dev);
        bg->setQUBO(b0, b1, Wb);
        bg->setPreference(sr::pnNumTrotters, mb);
        bg->setPreference(sr::Preference("rng", "mt19937"));
        bg->seed(11);
        bg->prepare();
        bg->randomizeSpin();
        for (int k = 0; k < 60; ++k)
            bg->annealOneStep(G[k], 50.f);
        bg->makeSolution();
        putVec(o, "bip_b0", b0.data, N0);
        putVec(o, "bip_b1", b1.data, N1);
        putVec(o, "bip_W", Wb.data, (size_t)N1 * N0);
        putVec(o, "bip_E", bg->get_E().data, (size_t)mb);
        sr::BitSetArray q0, q1;
        for (const sr::BitSetPair &p : bg->get_q()) {
            q0.pushBack(p.first);
            q1.pushBack(p.second);
        }
        putRows(o, "bip_q0", q0);
        putRows(o, "bip_q1", q1);
        sr::deleteInstance(bg);

        /* dense brute-force search */
        const int Nb = 12;
        sr::MatrixType<float> Wbf = symmetricW(Nb, 42);
        auto *bf = sr::newDenseGraphBFSearcher<float>();
        bf->assignDevice(*dev);
        bf->setQUBO(Wbf);
        bf->prepare();
        sr::PackedBitSet xEnd = 0;
        int ranges = 0;
        while (!bf->searchRange(&xEnd))
            ++ranges;
        bf->makeSolution();
        putVec(o, "bf_W", Wbf.data, (size_t)Nb * Nb);
        putVec(o, "bf_E", bf->get_E().data, (size_t)bf->get_E().size);
        putRows(o, "bf_x", bf->get_x());
        const double xe = (double)xEnd;
        putVec(o, "bf_x_end", &xe, 1);
        sr::deleteInstance(bf);

        /* dense formulas: E of every x in a batch */
        auto *fm = sr::newDenseGraphFormulas<float>();
        fm->assignDevice(*dev);
        sr::MatrixType<float> X(5, N);
        Lcg gx{5};
        for (int r = 0; r < 5; ++r)
            for (int i = 0; i < N; ++i)
                X(r, i) = (float)(gx.next() > 0 ? 1 : 0);
        sr::VectorType<float> Ex;
        fm->calculate_E(&Ex, W, X);
        putVec(o, "formulas_x", X.data, (size_t)5 * N);
        putVec(o, "formulas_E", Ex.data, (size_t)Ex.size);
        sr::deleteInstance(fm);

        /* errors are std::runtime_error, as the reference throws */
        bool threw = false;
        try {
            auto *bad = sr::newDenseGraphAnnealer<float>();
            try {
                bad->annealOneStep(1.f, 1.f); /* before setQUBO / prepare */
            } catch (const std::runtime_error &) {
                threw = true;
            }
            sr::deleteInstance(bad);
        } catch (...) {
        }
        const int t = threw ? 1 : 0;
        putVec(o, "state_error_thrown", &t, 1);
        sr::deleteInstance(dev);
    } catch (const std::exception &e) {
        fprintf(stderr, "cpp_api_run: %s\n", e.what());
        return 1;
    }
    o += "\"ok\": 1}";
    FILE *f = fopen(argv[1], "w");
    if (!f)
        return 1;
    fputs(o.c_str(), f);
    fclose(f);
    return 0;
}
