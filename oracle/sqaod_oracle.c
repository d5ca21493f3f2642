/*
 * sqaod CPU oracle -- TEST INFRASTRUCTURE ONLY (see sqaod_oracle.h).
 * Restates sqaod.cpu; the product library never links or loads this file.
 *
 * Build: gcc -O2 -msse2 -ffp-contract=off -fopenmp -fPIC -shared
 * (-ffp-contract=off: the reference CPU build targets SSE2 without FMA, so no
 * multiply-add is ever fused; this keeps the restatement's roundings equal.)
 */
#include "sqaod_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* MT19937: sqaodc/common/Random.cpp:64-168 (Matsumoto-Nishimura 2002)  */
/* ------------------------------------------------------------------ */
#define MT_N 624
#define MT_M 397

void sqo_mt_seed(sqo_mt *r, unsigned long long s) {
    /* Random.cpp:70-83: init_genrand with the low 32 bits of the seed */
    r->mt[0] = (uint32_t)(s & 0xffffffffull);
    for (int i = 1; i < MT_N; ++i)
        r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
    r->mti = MT_N;
}

static void mt_twist(sqo_mt *r) {
    /* Random.cpp:120-138: regenerate all 624 words */
    uint32_t *mt = r->mt;
    for (int k = 0; k < MT_N; ++k) {
        uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % MT_N] & 0x7fffffffu);
        uint32_t v = mt[(k + MT_M) % MT_N] ^ (y >> 1);
        if (y & 1u)
            v ^= 0x9908b0dfu;
        mt[k] = v;
    }
    r->mti = 0;
}

uint32_t sqo_mt_u32(sqo_mt *r) {
    if (r->mti >= MT_N)
        mt_twist(r);
    uint32_t y = r->mt[r->mti++];
    /* tempering, Random.cpp:142-146 */
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

uint32_t sqo_mt_randint(sqo_mt *r, int n) { return sqo_mt_u32(r) % (uint32_t)n; }

float sqo_mt_f32(sqo_mt *r) {
    /* Random.cpp:157-161: u32 converted to float (round to nearest even,
     * may give 2^32 -> 1.0f) then scaled by 2^-32 */
    return (float)sqo_mt_u32(r) * (float)(1. / 4294967296.);
}

double sqo_mt_f64(sqo_mt *r) {
    /* Random.cpp:165-168: 53-bit double from two draws, first draw high */
    uint32_t a = sqo_mt_u32(r) >> 5;
    uint32_t b = sqo_mt_u32(r) >> 6;
    return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

/* Random.cpp:157-168 conversions of raw words */
float sqo_word_to_f32(uint32_t w) { return (float)w * (float)(1. / 4294967296.); }
double sqo_words_to_f64(uint32_t w1, uint32_t w2) {
    uint32_t a = w1 >> 5, b = w2 >> 6;
    return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

void sqo_mt_stream(unsigned long long seed, unsigned long long skip, uint32_t *out, size_t n) {
    sqo_mt r;
    sqo_mt_seed(&r, seed);
    for (unsigned long long i = 0; i < skip; ++i)
        (void)sqo_mt_u32(&r);
    for (size_t i = 0; i < n; ++i)
        out[i] = sqo_mt_u32(&r);
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:  */
/* as easy as 1, 2, 3", SC'11): the counter-based stream of the rocm     */
/* device, restated with the device's counter layouts so that the CPU    */
/* annealers can consume the same stream.                               */
/* ------------------------------------------------------------------ */
void sqo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    /* ten rounds of two 32x32->64 multiplications (multipliers 0xD2511F53,
     * 0xCD9E8D57) with the Weyl key schedule (0x9E3779B9, 0xBB67AE85) */
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* dense attempt (y, k) of sweep s: counter (k, y, s_lo, s_hi), key = seed */
void sqo_philox_attempt(unsigned long long seed, unsigned long long sweep, uint32_t k, uint32_t y, uint32_t out[4]) {
    const uint32_t c[4] = {k, y, (uint32_t)sweep, (uint32_t)(sweep >> 32)};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    sqo_philox4x32_10(c, key, out);
}

/* naive attempt t of sweep s: counter (t_lo, t_hi, s_lo, s_hi ^ 0x20000000) */
void sqo_philox_naive(unsigned long long seed, unsigned long long sweep, unsigned long long t, uint32_t out[4]) {
    const uint32_t c[4] = {(uint32_t)t, (uint32_t)(t >> 32), (uint32_t)sweep,
                           (uint32_t)(sweep >> 32) ^ 0x20000000u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    sqo_philox4x32_10(c, key, out);
}

/* bipartite half-step of step index s, side 1 (q1 annealed) or 0:
 * counter (pos_lo, pos_hi, s_lo, s_hi ^ (side ? 0x80000000 : 0x40000000)) */
void sqo_philox_bg(unsigned long long seed, unsigned long long step, int side, unsigned long long pos,
                   uint32_t out[4]) {
    const uint32_t c[4] = {(uint32_t)pos, (uint32_t)(pos >> 32), (uint32_t)step,
                           (uint32_t)(step >> 32) ^ (side ? 0x80000000u : 0x40000000u)};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    sqo_philox4x32_10(c, key, out);
}

/* spins of an m x N block (row stride ld, elements of esize bytes, float or
 * double): q[y][x] = +1 if bit 0 of word x % 4 of the block at counter
 * (x / 4, y, salt, 0xffffffff) is set, else -1 */
void sqo_philox_spins(unsigned long long seed, uint32_t salt, int N, int m, void *Q, size_t esize, int ld) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int y = 0; y < m; ++y)
        for (int x0 = 0; x0 < N; x0 += 4) {
            const uint32_t c[4] = {(uint32_t)(x0 / 4), (uint32_t)y, salt, 0xffffffffu};
            uint32_t r[4];
            sqo_philox4x32_10(c, key, r);
            for (int b = 0; b < 4 && x0 + b < N; ++b) {
                double v = (r[b] & 1u) ? 1. : -1.;
                size_t i = (size_t)y * ld + x0 + b;
                if (esize == 4)
                    ((float *)Q)[i] = (float)v;
                else
                    ((double *)Q)[i] = v;
            }
        }
}

void sqo_philox_attempts(unsigned long long seed, unsigned long long sweep, int N, int m, uint32_t *X, void *U,
                         int dtype) {
    /* the device's per-sweep attempt arrays: X[y][k] = r0 % N, U[y][k] from
     * r1 (float) or (r1, r2) (double) */
    for (int y = 0; y < m; ++y)
        for (int k = 0; k < N; ++k) {
            uint32_t r[4];
            sqo_philox_attempt(seed, sweep, (uint32_t)k, (uint32_t)y, r);
            size_t i = (size_t)y * N + k;
            X[i] = r[0] % (uint32_t)N;
            if (dtype)
                ((double *)U)[i] = sqo_words_to_f64(r[1], r[2]);
            else
                ((float *)U)[i] = sqo_word_to_f32(r[1]);
        }
}

/* ------------------------------------------------------------------ */
/* dot accumulated in double, 16 lanes in a fixed order (the bipartite   */
/* half-step GEMM of the oracle; exact for integer data)                */
/* ------------------------------------------------------------------ */
#if defined(__x86_64__)
#include <immintrin.h>
__attribute__((target("avx2"))) static double dot64_f32_avx2(const float *a, const float *b, int n) {
    __m256d s[4] = {_mm256_setzero_pd(), _mm256_setzero_pd(), _mm256_setzero_pd(), _mm256_setzero_pd()};
    int i = 0;
    for (; i + 16 <= n; i += 16)
        for (int l = 0; l < 4; ++l)
            s[l] = _mm256_add_pd(s[l], _mm256_mul_pd(_mm256_cvtps_pd(_mm_loadu_ps(a + i + 4 * l)),
                                                     _mm256_cvtps_pd(_mm_loadu_ps(b + i + 4 * l))));
    double lane[16];
    for (int l = 0; l < 4; ++l)
        _mm256_storeu_pd(lane + 4 * l, s[l]);
    for (; i < n; ++i)
        lane[i & 15] += (double)a[i] * (double)b[i];
    double r = 0.;
    for (int l = 0; l < 16; ++l)
        r += lane[l];
    return r;
}
#endif
static double dot64_f32_plain(const float *a, const float *b, int n) {
    double lane[16] = {0.};
    int i = 0;
    for (; i + 16 <= n; i += 16)
        for (int l = 0; l < 16; ++l)
            lane[l] += (double)a[i + l] * (double)b[i + l];
    for (; i < n; ++i)
        lane[i & 15] += (double)a[i] * (double)b[i];
    double r = 0.;
    for (int l = 0; l < 16; ++l)
        r += lane[l];
    return r;
}
static double dot64_f32(const float *a, const float *b, int n) {
#if defined(__x86_64__)
    static int avx2 = -1;
    if (avx2 < 0)
        avx2 = __builtin_cpu_supports("avx2") ? 1 : 0;
    if (avx2)
        return dot64_f32_avx2(a, b, n);
#endif
    return dot64_f32_plain(a, b, n);
}
static double dot64_f64(const double *a, const double *b, int n) {
    double lane[16] = {0.};
    int i = 0;
    for (; i + 16 <= n; i += 16)
        for (int l = 0; l < 16; ++l)
            lane[l] += a[i + l] * b[i + l];
    for (; i < n; ++i)
        lane[i & 15] += a[i] * b[i];
    double r = 0.;
    for (int l = 0; l < 16; ++l)
        r += lane[l];
    return r;
}

/* ------------------------------------------------------------------ */
/* dot in the SSE2 order: sqaodc/cpu/Dot_SIMD.cpp:15-110                */
/* ------------------------------------------------------------------ */
typedef float v4f __attribute__((vector_size(16)));
typedef double v2d __attribute__((vector_size(16)));

float sqo_dot_f32(const float *v0, const float *v1, int n) {
    /* Blocks of 16 elements (zero padding read past n); within a block four
     * 4-lane products are added pairwise ((p0+p1)+(p2+p3)) and accumulated
     * per lane; the lanes are combined as (l0+l2)+(l1+l3).  Dot_SIMD.cpp:60-110. */
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    int nBlocks = (n + 15) / 16;
    for (int b = 0; b < nBlocks; ++b) {
        v4f p[4];
        for (int k = 0; k < 4; ++k) {
            v4f a, c;
            for (int l = 0; l < 4; ++l) {
                int i = b * 16 + k * 4 + l;
                a[l] = i < n ? v0[i] : 0.f;
                c[l] = i < n ? v1[i] : 0.f;
            }
            p[k] = a * c;
        }
        acc = acc + ((p[0] + p[1]) + (p[2] + p[3]));
    }
    float lo = acc[0] + acc[2];
    float hi = acc[1] + acc[3];
    return lo + hi;
}

double sqo_dot_f64(const double *v0, const double *v1, int n) {
    /* Blocks of 8 elements as four 2-lane products, ((p0+p1)+(p2+p3)) per
     * lane; lanes combined as l0 + l1.  Dot_SIMD.cpp:15-55. */
    v2d acc = {0., 0.};
    int nBlocks = (n + 7) / 8;
    for (int b = 0; b < nBlocks; ++b) {
        v2d p[4];
        for (int k = 0; k < 4; ++k) {
            v2d a, c;
            for (int l = 0; l < 2; ++l) {
                int i = b * 8 + k * 2 + l;
                a[l] = i < n ? v0[i] : 0.;
                c[l] = i < n ? v1[i] : 0.;
            }
            p[k] = a * c;
        }
        acc = acc + ((p[0] + p[1]) + (p[2] + p[3]));
    }
    return acc[0] + acc[1];
}

/* ------------------------------------------------------------------ */
/* dense formulas in double: SharedFormulas.cpp:23-83                   */
/* ------------------------------------------------------------------ */
void sqo_dg_energy_q_f64acc(int N, int m, const double *h, const double *J, double c,
                            const int8_t *q, double *E) {
    /* E_y = -c - h.q_y - q_y^T J q_y   (SharedFormulas.cpp:70-83) */
#pragma omp parallel for schedule(static) if ((size_t)m * N * N > ((size_t)1 << 18))
    for (int y = 0; y < m; ++y) {
        const int8_t *qy = q + (size_t)y * N;
        double hq = 0., qJq = 0.;
        for (int i = 0; i < N; ++i) {
            double s = 0.;
            const double *Ji = J + (size_t)i * N;
            for (int j = 0; j < N; ++j)
                s += Ji[j] * (double)qy[j];
            qJq += (double)qy[i] * s;
            hq += h[i] * (double)qy[i];
        }
        E[y] = -c - hq - qJq;
    }
}

void sqo_dg_energy_x_f64acc(int N, int nx, const double *W, const int8_t *x, double *E) {
    /* E = x^T W x  (SharedFormulas.cpp:23-32) */
#pragma omp parallel for schedule(static) if ((size_t)nx * N * N > ((size_t)1 << 18))
    for (int k = 0; k < nx; ++k) {
        const int8_t *xk = x + (size_t)k * N;
        double e = 0.;
        for (int i = 0; i < N; ++i) {
            if (!xk[i])
                continue;
            const double *Wi = W + (size_t)i * N;
            for (int j = 0; j < N; ++j)
                e += Wi[j] * (double)xk[j];
        }
        E[k] = e;
    }
}

/* ------------------------------------------------------------------ */
/* typed bodies                                                         */
/* ------------------------------------------------------------------ */
#define REAL float
#define SFX(name) name##_f32
#define EXPR(v) expf(v)
#define TANHR(v) tanhf(v)
#define LOGR(v) logf(v)
#define DOTR(a, b, n) sqo_dot_f32(a, b, n)
#define RANDR(r) sqo_mt_f32(r)
#define PHILOX_U0(p) sqo_word_to_f32((p)[0])
#define DOT64(a, b, n) dot64_f32(a, b, n)
#define PAD 16
#include "oracle_typed.inc"
#undef REAL
#undef SFX
#undef EXPR
#undef TANHR
#undef LOGR
#undef DOTR
#undef RANDR
#undef PHILOX_U0
#undef DOT64
#undef PAD

#define REAL double
#define SFX(name) name##_f64
#define EXPR(v) exp(v)
#define TANHR(v) tanh(v)
#define LOGR(v) log(v)
#define DOTR(a, b, n) sqo_dot_f64(a, b, n)
#define RANDR(r) sqo_mt_f64(r)
#define PHILOX_U0(p) sqo_words_to_f64((p)[0], (p)[1])
#define DOT64(a, b, n) dot64_f64(a, b, n)
#define PAD 8
#include "oracle_typed.inc"
#undef REAL
#undef SFX
#undef EXPR
#undef TANHR
#undef LOGR
#undef DOTR
#undef RANDR
#undef PHILOX_U0
#undef DOT64
#undef PAD

/* ------------------------------------------------------------------ */
/* dtype dispatch                                                       */
/* ------------------------------------------------------------------ */
struct sqo_dg {
    int dtype;
    void *impl;
};

sqo_dg *sqo_dg_new(int dtype, int nWorkers) {
    sqo_dg *a = (sqo_dg *)calloc(1, sizeof(sqo_dg));
    a->dtype = dtype;
    a->impl = dtype ? (void *)dg_new_f64(nWorkers) : (void *)dg_new_f32(nWorkers);
    return a;
}
void sqo_dg_free(sqo_dg *a) {
    if (!a)
        return;
    if (a->dtype) dg_free_f64(a->impl); else dg_free_f32(a->impl);
    free(a);
}
#define DG_DISPATCH(call_f32, call_f64) \
    do { if (a->dtype) { call_f64; } else { call_f32; } } while (0)

void sqo_dg_seed(sqo_dg *a, unsigned long long s) {
    DG_DISPATCH(dg_seed_f32(a->impl, s), dg_seed_f64(a->impl, s));
}
void sqo_dg_set_qubo(sqo_dg *a, int N, const void *W, int om) {
    DG_DISPATCH(dg_set_qubo_f32(a->impl, N, (const float *)W, om),
                dg_set_qubo_f64(a->impl, N, (const double *)W, om));
}
void sqo_dg_set_hamiltonian(sqo_dg *a, int N, const void *h, const void *J, double c) {
    DG_DISPATCH(dg_set_hamiltonian_f32(a->impl, N, (const float *)h, (const float *)J, (float)c),
                dg_set_hamiltonian_f64(a->impl, N, (const double *)h, (const double *)J, c));
}
void sqo_dg_get_hamiltonian(const sqo_dg *a, void *h, void *J, double *c) {
    DG_DISPATCH(dg_get_hamiltonian_f32(a->impl, (float *)h, (float *)J, c),
                dg_get_hamiltonian_f64(a->impl, (double *)h, (double *)J, c));
}
void sqo_dg_set_m(sqo_dg *a, int m) { DG_DISPATCH(dg_set_m_f32(a->impl, m), dg_set_m_f64(a->impl, m)); }
int sqo_dg_get_m(const sqo_dg *a) { return a->dtype ? dg_get_m_f64(a->impl) : dg_get_m_f32(a->impl); }
void sqo_dg_set_algorithm(sqo_dg *a, int algo) {
    DG_DISPATCH(dg_set_algorithm_f32(a->impl, algo), dg_set_algorithm_f64(a->impl, algo));
}
int sqo_dg_prepare(sqo_dg *a) { return a->dtype ? dg_prepare_f64(a->impl) : dg_prepare_f32(a->impl); }
void sqo_dg_randomize_spin(sqo_dg *a) {
    DG_DISPATCH(dg_randomize_spin_f32(a->impl), dg_randomize_spin_f64(a->impl));
}
void sqo_dg_set_q(sqo_dg *a, const int8_t *q) { DG_DISPATCH(dg_set_q_f32(a->impl, q), dg_set_q_f64(a->impl, q)); }
void sqo_dg_get_q(const sqo_dg *a, int8_t *q) { DG_DISPATCH(dg_get_q_f32(a->impl, q), dg_get_q_f64(a->impl, q)); }
void sqo_dg_anneal_one_step(sqo_dg *a, double G, double beta) {
    DG_DISPATCH(dg_anneal_one_step_f32(a->impl, (float)G, (float)beta),
                dg_anneal_one_step_f64(a->impl, G, beta));
}
void sqo_dg_calculate_E(sqo_dg *a, void *E) {
    DG_DISPATCH(dg_calculate_E_f32(a->impl, (float *)E), dg_calculate_E_f64(a->impl, (double *)E));
}
double sqo_dg_get_system_E(sqo_dg *a, double G, double beta) {
    return a->dtype ? dg_get_system_E_f64(a->impl, G, beta)
                    : (double)dg_get_system_E_f32(a->impl, (float)G, (float)beta);
}
void sqo_dg_set_stream(sqo_dg *a, int kind) { DG_DISPATCH(dg_set_stream_f32(a->impl, kind), dg_set_stream_f64(a->impl, kind)); }
void sqo_dg_set_local_fields(sqo_dg *a, int on) { DG_DISPATCH(dg_set_local_f32(a->impl, on), dg_set_local_f64(a->impl, on)); }
void sqo_dg_anneal_one_step_injected(sqo_dg *a, double G, double beta, const uint32_t *Y, const uint32_t *X,
                                     const void *U) {
    DG_DISPATCH(dg_anneal_one_step_injected_f32(a->impl, (float)G, (float)beta, Y, X, (const float *)U),
                dg_anneal_one_step_injected_f64(a->impl, G, beta, Y, X, (const double *)U));
}
long long sqo_dg_last_flips(const sqo_dg *a) {
    return a->dtype ? dg_last_flips_f64(a->impl) : dg_last_flips_f32(a->impl);
}

/* ---- bipartite dispatch ---- */
struct sqo_bg {
    int dtype;
    void *impl;
};

sqo_bg *sqo_bg_new(int dtype, int nWorkers) {
    sqo_bg *a = (sqo_bg *)calloc(1, sizeof(sqo_bg));
    a->dtype = dtype;
    a->impl = dtype ? (void *)bg_new_f64(nWorkers) : (void *)bg_new_f32(nWorkers);
    return a;
}
void sqo_bg_free(sqo_bg *a) {
    if (!a)
        return;
    if (a->dtype) bg_free_f64(a->impl); else bg_free_f32(a->impl);
    free(a);
}
void sqo_bg_seed(sqo_bg *a, unsigned long long s) {
    DG_DISPATCH(bg_seed_f32(a->impl, s), bg_seed_f64(a->impl, s));
}
void sqo_bg_set_qubo(sqo_bg *a, int N0, int N1, const void *b0, const void *b1, const void *W, int om) {
    DG_DISPATCH(bg_set_qubo_f32(a->impl, N0, N1, (const float *)b0, (const float *)b1, (const float *)W, om),
                bg_set_qubo_f64(a->impl, N0, N1, (const double *)b0, (const double *)b1, (const double *)W, om));
}
void sqo_bg_set_hamiltonian(sqo_bg *a, int N0, int N1, const void *h0, const void *h1, const void *J, double c) {
    DG_DISPATCH(bg_set_hamiltonian_f32(a->impl, N0, N1, (const float *)h0, (const float *)h1, (const float *)J, (float)c),
                bg_set_hamiltonian_f64(a->impl, N0, N1, (const double *)h0, (const double *)h1, (const double *)J, c));
}
void sqo_bg_get_hamiltonian(const sqo_bg *a, void *h0, void *h1, void *J, double *c) {
    DG_DISPATCH(bg_get_hamiltonian_f32(a->impl, (float *)h0, (float *)h1, (float *)J, c),
                bg_get_hamiltonian_f64(a->impl, (double *)h0, (double *)h1, (double *)J, c));
}
void sqo_bg_set_m(sqo_bg *a, int m) { DG_DISPATCH(bg_set_m_f32(a->impl, m), bg_set_m_f64(a->impl, m)); }
int sqo_bg_get_m(const sqo_bg *a) { return a->dtype ? bg_get_m_f64(a->impl) : bg_get_m_f32(a->impl); }
void sqo_bg_set_algorithm(sqo_bg *a, int algo) {
    DG_DISPATCH(bg_set_algorithm_f32(a->impl, algo), bg_set_algorithm_f64(a->impl, algo));
}
int sqo_bg_prepare(sqo_bg *a) { return a->dtype ? bg_prepare_f64(a->impl) : bg_prepare_f32(a->impl); }
void sqo_bg_randomize_spin(sqo_bg *a) {
    DG_DISPATCH(bg_randomize_spin_f32(a->impl), bg_randomize_spin_f64(a->impl));
}
void sqo_bg_set_q(sqo_bg *a, const int8_t *q0, const int8_t *q1) {
    DG_DISPATCH(bg_set_q_f32(a->impl, q0, q1), bg_set_q_f64(a->impl, q0, q1));
}
void sqo_bg_get_q(const sqo_bg *a, int8_t *q0, int8_t *q1) {
    DG_DISPATCH(bg_get_q_f32(a->impl, q0, q1), bg_get_q_f64(a->impl, q0, q1));
}
void sqo_bg_anneal_one_step(sqo_bg *a, double G, double beta) {
    DG_DISPATCH(bg_anneal_one_step_f32(a->impl, (float)G, (float)beta),
                bg_anneal_one_step_f64(a->impl, G, beta));
}
void sqo_bg_calculate_E(sqo_bg *a, void *E) {
    DG_DISPATCH(bg_calculate_E_f32(a->impl, (float *)E), bg_calculate_E_f64(a->impl, (double *)E));
}
void sqo_bg_set_stream(sqo_bg *a, int kind) { DG_DISPATCH(bg_set_stream_f32(a->impl, kind), bg_set_stream_f64(a->impl, kind)); }
void sqo_bg_set_local_fields(sqo_bg *a, int on) { DG_DISPATCH(bg_set_local_f32(a->impl, on), bg_set_local_f64(a->impl, on)); }
double sqo_bg_get_system_E(sqo_bg *a, double G, double beta) {
    return a->dtype ? bg_get_system_E_f64(a->impl, G, beta)
                    : (double)bg_get_system_E_f32(a->impl, (float)G, (float)beta);
}

/* ------------------------------------------------------------------ */
/* dense brute-force search: CPUDenseGraphBFSearcher.cpp:64-198 and    */
/* CPUDenseGraphBatchSearch.cpp:18-50, single-worker semantics          */
/* ------------------------------------------------------------------ */
typedef struct {
    double Emin;
    int n;
    unsigned long long *xs;
} bf_range_t;

static inline double bf_round(int dtype, double e) { return dtype ? e : (double)(float)e; }

int sqo_dg_bf_search(int dtype, int N, const double *W, int om, int tileSize,
                     double *EminOut, unsigned long long *xs) {
    /* E(x) = x^T W x with bit (N-1-pos) of the packed x holding x[pos]
     * (Common.cpp:79-86).  x is scanned in ascending order; each contiguous
     * range keeps (Emin, first tileSize ties) exactly as one BatchSearch does,
     * and the ranges are merged in order, which reproduces the one-worker
     * result.  Energies are exact in double (symmetric part of W, incremental
     * over binary counting) and rounded to the solver's real type before
     * comparison. */
    if (N < 1 || N > 40)
        return -1;
    unsigned long long xMax = 1ull << N;
    if ((unsigned long long)tileSize > xMax)
        tileSize = (int)xMax; /* :70-73 */
    double sgn = (om == SQO_MAXIMIZE) ? -1. : 1.;
    double *S = (double *)malloc((size_t)N * N * sizeof(double));
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            S[(size_t)i * N + j] = sgn * 0.5 * (W[(size_t)i * N + j] + W[(size_t)j * N + i]);
    int nRanges = 1;
    while (nRanges < 512 && ((xMax / (unsigned long long)(nRanges * 2)) >= 1024))
        nRanges *= 2;
    unsigned long long span = xMax / (unsigned long long)nRanges;
    bf_range_t *rs = (bf_range_t *)calloc((size_t)nRanges, sizeof(bf_range_t));
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < nRanges; ++r) {
        bf_range_t *R = &rs[r];
        R->Emin = FLT_MAX; /* :20 */
        R->n = 0;
        R->xs = (unsigned long long *)malloc((size_t)tileSize * sizeof(unsigned long long));
        double *f = (double *)calloc((size_t)N, sizeof(double));
        int *b = (int *)calloc((size_t)N, sizeof(int));
        unsigned long long x0 = span * (unsigned long long)r;
        for (int pos = 0; pos < N; ++pos)
            b[pos] = (int)((x0 >> (N - 1 - pos)) & 1ull);
        double E = 0.;
        for (int i = 0; i < N; ++i) {
            double s = 0.;
            for (int j = 0; j < N; ++j)
                s += S[(size_t)i * N + j] * (double)b[j];
            f[i] = s;
            E += (double)b[i] * s;
        }
        for (unsigned long long x = x0; x < x0 + span; ++x) {
            if (x != x0) {
                /* x-1 -> x: bits 0..t change, t = trailing ones of x-1 */
                unsigned long long prev = x - 1;
                int t = 0;
                while ((prev >> t) & 1ull)
                    ++t;
                for (int bit = 0; bit <= t; ++bit) {
                    int k = N - 1 - bit;
                    const double *Sk = S + (size_t)k * N;
                    if (b[k]) {
                        E += -2. * f[k] + Sk[k];
                        for (int i = 0; i < N; ++i) f[i] -= Sk[i];
                        b[k] = 0;
                    } else {
                        E += 2. * f[k] + Sk[k];
                        for (int i = 0; i < N; ++i) f[i] += Sk[i];
                        b[k] = 1;
                    }
                }
            }
            double Er = bf_round(dtype, E);
            if (Er > R->Emin)
                continue;
            if (Er == R->Emin) {
                if (R->n < tileSize)
                    R->xs[R->n++] = x;
            } else {
                R->Emin = Er;
                R->n = 0;
                R->xs[R->n++] = x;
            }
        }
        free(f);
        free(b);
    }
    double Emin = FLT_MAX;
    int n = 0;
    for (int r = 0; r < nRanges; ++r) {
        bf_range_t *R = &rs[r];
        if (R->Emin < Emin) {
            Emin = R->Emin;
            n = 0;
            for (int k = 0; k < R->n && n < tileSize; ++k)
                xs[n++] = R->xs[k];
        } else if (R->Emin == Emin) {
            for (int k = 0; k < R->n && n < tileSize; ++k)
                xs[n++] = R->xs[k];
        }
        free(R->xs);
    }
    free(rs);
    free(S);
    /* calculate_E reports -Emin when maximizing, :92-101 */
    *EminOut = (om == SQO_MAXIMIZE) ? -Emin : Emin;
    return n;
}

/* ------------------------------------------------------------------ */
/* bipartite brute-force search: CPUBipartiteGraphBFSearcher.cpp:75-225 */
/* and CPUBipartiteGraphBatchSearch.cpp:29-62, single-worker semantics  */
/* ------------------------------------------------------------------ */
int sqo_bg_bf_search(int dtype, int N0, int N1, const double *b0, const double *b1, const double *W, int om,
                     int tileSize0, int tileSize1, double *EminOut, unsigned long long *xs0,
                     unsigned long long *xs1) {
    /* E(x0, x1) = b0.x0 + b1.x1 + x1^T W x0 (calculate_E_2d, SharedFormulas.cpp:123-144; W is
     * N1 x N0, negated with b0, b1 when maximizing, :39-47), bit (N-1-pos) of a packed x holding
     * x[pos] (Common.cpp:79-86).  One worker visits x0 tiles in order, and inside each the x1
     * tiles in order (searchRangeSingleThread, :166-198); a tile pair is scanned x1-major
     * (BatchSearch :47-60), so pairs are met in the order (x0 / T0, x1, x0 % T0).  Strictly lower
     * E restarts the list, equal E appends while it holds fewer than N0 + N1 pairs (:45, :52).
     * Energies are exact in double and rounded to the solver's real type before comparison. */
    if (N0 < 1 || N1 < 1 || N0 > 30 || N1 > 30 || N0 + N1 > 40)
        return -1;
    const unsigned long long x0max = 1ull << N0, x1max = 1ull << N1;
    unsigned long long T0 = (unsigned long long)tileSize0;
    if (T0 > x0max)
        T0 = x0max; /* prepare, :78-85 */
    (void)tileSize1; /* x1 tiles are visited in order, so they do not change the pair order */
    const double sgn = (om == SQO_MAXIMIZE) ? -1. : 1.;
    const int maxN = N0 + N1;
    double Emin = FLT_MAX;
    int n = 0;
    for (unsigned long long t0 = 0; t0 < x0max; t0 += T0) {
        for (unsigned long long x1 = 0; x1 < x1max; ++x1) {
            double bx1 = 0.;
            for (int i = 0; i < N1; ++i)
                if ((x1 >> (N1 - 1 - i)) & 1ull)
                    bx1 += sgn * b1[i];
            for (unsigned long long x0 = t0; x0 < t0 + T0; ++x0) {
                double E = bx1;
                for (int j = 0; j < N0; ++j)
                    if ((x0 >> (N0 - 1 - j)) & 1ull)
                        E += sgn * b0[j];
                for (int i = 0; i < N1; ++i) {
                    if (!((x1 >> (N1 - 1 - i)) & 1ull))
                        continue;
                    const double *Wi = W + (size_t)i * N0;
                    for (int j = 0; j < N0; ++j)
                        if ((x0 >> (N0 - 1 - j)) & 1ull)
                            E += sgn * Wi[j];
                }
                const double Er = bf_round(dtype, E);
                if (Er > Emin)
                    continue;
                if (Er == Emin) {
                    if (n < maxN) {
                        xs0[n] = x0;
                        xs1[n] = x1;
                        ++n;
                    }
                } else {
                    Emin = Er;
                    n = 0;
                    xs0[n] = x0;
                    xs1[n] = x1;
                    ++n;
                }
            }
        }
    }
    *EminOut = (om == SQO_MAXIMIZE) ? -Emin : Emin; /* calculate_E, :114-123 */
    return n;
}

/* ---- bipartite formulas in double ---- */
void sqo_bg_hamiltonian_f64(int N0, int N1, const double *b0, const double *b1,
                            const double *W, double *h0, double *h1, double *J, double *c) {
    bg_hamiltonian_f64(N0, N1, b0, b1, W, h0, h1, J, c); /* SharedFormulas.cpp:147-167 */
}

void sqo_bg_energy_q_f64acc(int N0, int N1, int m, const double *h0, const double *h1,
                            const double *J, double c, const int8_t *q0, const int8_t *q1,
                            double *E) {
    /* E = -(q1^T J q0) - h0.q0 - h1.q1 - c   (SharedFormulas.cpp:184-202) */
#pragma omp parallel for schedule(static)
    for (int im = 0; im < m; ++im) {
        const int8_t *a = q0 + (size_t)im * N0, *b = q1 + (size_t)im * N1;
        double e = 0., s0 = 0., s1 = 0.;
        for (int i = 0; i < N1; ++i) {
            double s = 0.;
            for (int j = 0; j < N0; ++j)
                s += J[(size_t)i * N0 + j] * (double)a[j];
            e += s * (double)b[i];
        }
        for (int j = 0; j < N0; ++j) s0 += h0[j] * (double)a[j];
        for (int i = 0; i < N1; ++i) s1 += h1[i] * (double)b[i];
        E[im] = -e - s0 - s1 - c;
    }
}

/* ---- host expf checksum, the counterpart of sqaod_rocm_expf_checksum ---- */
static inline unsigned long long sqo_mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

unsigned long long sqo_expf_checksum(unsigned long long lo, unsigned long long hi) {
    /* sum over b in [lo, hi) of splitmix64((b << 32) | bits(expf(b))) with the
     * host C library's expf (what CPUDenseGraphAnnealer.cpp:262 calls) */
    unsigned long long acc = 0;
#pragma omp parallel for schedule(static, 1 << 20) reduction(+ : acc)
    for (long long b = (long long)lo; b < (long long)hi; ++b) {
        uint32_t bits = (uint32_t)b;
        float x;
        memcpy(&x, &bits, 4);
        float r = expf(x);
        uint32_t rb;
        memcpy(&rb, &r, 4);
        if (r != r)
            rb = 0x7fc00000u;
        acc += sqo_mix64(((unsigned long long)b << 32) | rb);
    }
    return acc;
}
