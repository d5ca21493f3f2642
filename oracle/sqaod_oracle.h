/*
 * sqaod CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the sqaod.cpu semantics for the simulated-quantum-
 * annealing hot path (dense and bipartite coloring SQA, SA, dense brute-force
 * search, MT19937, Hamiltonian and energy formulas).  It exists so that tests,
 * the smoke check and the CPU-baseline leg of bench.py can check and time the
 * HIP implementation; the product library never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the sqaod source tree (sqaodc/...).
 *
 * Pinning (see tests/test_oracle_*.py):
 *  - MT19937 / randInt / randomf32 / randomf64: against the reference's own
 *    Random.cpp compiled from source into oracle/_ref, against the textbook
 *    known answer (10000th output of the default-seeded generator is
 *    4123659995) and the closed-form conversions.
 *  - dot (SSE2 order): against an exact (rational/fp64) sum for integer data
 *    and a numpy restatement of the 16-element block order for float data.
 *  - Hamiltonian / energies: closed forms (E(x) = x^T W x with x = (q+1)/2),
 *    the identity -sum(h) + sum(J) + c = 0 and the reference's test fixtures.
 *  - annealing trajectories: invariants (tracked vs recomputed energy,
 *    trivial problems reach E = 0) and brute force on tiny inputs; the
 *    trajectory restatement itself is "parity pinned by invariants only"
 *    because no reference test pins a trajectory.
 *  - brute-force search: direct O(N^2) evaluation at small N and the
 *    +-ones known answers.
 */
#ifndef SQAOD_ORACLE_H
#define SQAOD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MT19937 (sqaodc/common/Random.cpp:64-168) ---- */
typedef struct {
    uint32_t mt[624];
    int mti;
} sqo_mt;

void sqo_mt_seed(sqo_mt *r, unsigned long long s);        /* Random.cpp:70-83 */
uint32_t sqo_mt_u32(sqo_mt *r);                            /* Random.cpp:115-149 */
uint32_t sqo_mt_randint(sqo_mt *r, int n);                 /* Random.cpp:151-153 */
float sqo_mt_f32(sqo_mt *r);                               /* Random.cpp:157-161 */
double sqo_mt_f64(sqo_mt *r);                              /* Random.cpp:165-168 */
/* n consecutive tempered outputs after skipping `skip` outputs. */
void sqo_mt_stream(unsigned long long seed, unsigned long long skip, uint32_t *out, size_t n);

float sqo_word_to_f32(uint32_t w);                          /* Random.cpp:157-161 */
double sqo_words_to_f64(uint32_t w1, uint32_t w2);          /* Random.cpp:165-168 */

/* ---- Philox4x32-10 (Salmon et al., SC'11) in the rocm device's layouts ---- */
void sqo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void sqo_philox_attempt(unsigned long long seed, unsigned long long sweep, uint32_t k, uint32_t y, uint32_t out[4]);
void sqo_philox_naive(unsigned long long seed, unsigned long long sweep, unsigned long long t, uint32_t out[4]);
void sqo_philox_bg(unsigned long long seed, unsigned long long step, int side, unsigned long long pos,
                   uint32_t out[4]);
void sqo_philox_spins(unsigned long long seed, uint32_t salt, int N, int m, void *Q, size_t esize, int ld);
/* X[y][k], U[y][k] (dtype 0 float, 1 double) of one dense sweep */
void sqo_philox_attempts(unsigned long long seed, unsigned long long sweep, int N, int m, uint32_t *X, void *U,
                         int dtype);

/* ---- SIMD dot order (sqaodc/cpu/Dot_SIMD.cpp:15-110) ---- */
float sqo_dot_f32(const float *v0, const float *v1, int n);    /* dot_sse2(float), :60-110 */
double sqo_dot_f64(const double *v0, const double *v1, int n); /* dot_sse2(double), :15-55 */

/* ---- dense formulas (sqaodc/cpu/SharedFormulas.cpp) ---- */
/* W, J: N x N row-major, unpadded. */
void sqo_dg_hamiltonian_f32(int N, const float *W, float *h, float *J, float *c);   /* :36-55 */
void sqo_dg_hamiltonian_f64(int N, const double *W, double *h, double *J, double *c);
/* E[y] = -c - h.q_y - q_y^T J q_y, accumulated in double (the FP64 recompute). */
void sqo_dg_energy_q_f64acc(int N, int m, const double *h, const double *J, double c,
                            const int8_t *q, double *E);                            /* :70-83 */
/* E[k] = x_k^T W x_k, accumulated in double. */
void sqo_dg_energy_x_f64acc(int N, int nx, const double *W, const int8_t *x, double *E); /* :23-32 */

/* ---- dense annealer (sqaodc/cpu/CPUDenseGraphAnnealer.cpp) ---- */
enum { SQO_ALGO_NAIVE = 2, SQO_ALGO_COLORING = 3, SQO_ALGO_SA_NAIVE = 6 }; /* Preference.h:8-17 */
enum { SQO_MINIMIZE = 0, SQO_MAXIMIZE = 1 };                                /* Solver.h:9-14 */
/* attempt streams: sqaod.cpu's per-worker MT19937 (default), the rocm
 * device's Philox layout, or arrays injected per step */
enum { SQO_STREAM_MT = 0, SQO_STREAM_PHILOX = 1, SQO_STREAM_INJECTED = 2 };

typedef struct sqo_dg sqo_dg;
/* dtype: 0 = float, 1 = double.  nWorkers: size of the per-worker RNG array
 * (CPUDenseGraphAnnealer.cpp:18-20); 1 selects the serial coloring variant. */
sqo_dg *sqo_dg_new(int dtype, int nWorkers);
void sqo_dg_free(sqo_dg *a);
void sqo_dg_seed(sqo_dg *a, unsigned long long seed);                 /* :29-33 */
void sqo_dg_set_qubo(sqo_dg *a, int N, const void *W, int om);        /* :52-69 */
void sqo_dg_set_hamiltonian(sqo_dg *a, int N, const void *h, const void *J, double c); /* :72-84 */
void sqo_dg_set_m(sqo_dg *a, int m);
int sqo_dg_get_m(const sqo_dg *a);
void sqo_dg_set_algorithm(sqo_dg *a, int algo);
int sqo_dg_prepare(sqo_dg *a);                                        /* :164-195; returns algo */
void sqo_dg_randomize_spin(sqo_dg *a);                                /* :152-161 */
void sqo_dg_set_q(sqo_dg *a, const int8_t *q);                        /* m x N, :123-134 */
void sqo_dg_get_q(const sqo_dg *a, int8_t *q);                        /* m x N, :216-227 */
void sqo_dg_get_hamiltonian(const sqo_dg *a, void *h, void *J, double *c);
void sqo_dg_anneal_one_step(sqo_dg *a, double G, double beta);        /* :268-371 */
void sqo_dg_calculate_E(sqo_dg *a, void *E);                          /* :206-213 */
double sqo_dg_get_system_E(sqo_dg *a, double G, double beta);         /* :230-247 */
/* number of accepted flips in the last anneal_one_step (diagnostic) */
long long sqo_dg_last_flips(const sqo_dg *a);
void sqo_dg_set_stream(sqo_dg *a, int kind);
/* incremental local fields instead of the per-attempt dot: the same results
 * for integer-valued problems (every sum exact), much faster at large N */
void sqo_dg_set_local_fields(sqo_dg *a, int on);
/* one step consuming the caller's stream: coloring/SA take X, U of m x N in
 * (slice, attempt) order; naive takes Y, X, U of N m entries in serial order */
void sqo_dg_anneal_one_step_injected(sqo_dg *a, double G, double beta, const uint32_t *Y, const uint32_t *X,
                                     const void *U);

/* ---- bipartite annealer (sqaodc/cpu/CPUBipartiteGraphAnnealer.cpp) ---- */
typedef struct sqo_bg sqo_bg;
sqo_bg *sqo_bg_new(int dtype, int nWorkers);
void sqo_bg_free(sqo_bg *a);
void sqo_bg_seed(sqo_bg *a, unsigned long long seed);                 /* :31-36 */
/* b0[N0], b1[N1], W[N1][N0] */
void sqo_bg_set_qubo(sqo_bg *a, int N0, int N1, const void *b0, const void *b1,
                     const void *W, int om);                          /* :54-75 */
void sqo_bg_set_hamiltonian(sqo_bg *a, int N0, int N1, const void *h0, const void *h1,
                            const void *J, double c);                 /* :77-92 */
void sqo_bg_get_hamiltonian(const sqo_bg *a, void *h0, void *h1, void *J, double *c);
void sqo_bg_set_m(sqo_bg *a, int m);
int sqo_bg_get_m(const sqo_bg *a);
void sqo_bg_set_algorithm(sqo_bg *a, int algo);
int sqo_bg_prepare(sqo_bg *a);                                        /* :224-258 */
void sqo_bg_randomize_spin(sqo_bg *a);                                /* :172-213 */
void sqo_bg_set_q(sqo_bg *a, const int8_t *q0, const int8_t *q1);     /* m x N0, m x N1 */
void sqo_bg_get_q(const sqo_bg *a, int8_t *q0, int8_t *q1);
void sqo_bg_anneal_one_step(sqo_bg *a, double G, double beta);        /* :288-542 */
void sqo_bg_calculate_E(sqo_bg *a, void *E);                          /* :215-222 */
double sqo_bg_get_system_E(sqo_bg *a, double G, double beta);         /* :268-286 */
void sqo_bg_set_stream(sqo_bg *a, int kind);                          /* MT (per worker) or Philox */
void sqo_bg_set_local_fields(sqo_bg *a, int on);                      /* incremental dEmat (integer data) */
void sqo_bg_hamiltonian_f64(int N0, int N1, const double *b0, const double *b1,
                            const double *W, double *h0, double *h1, double *J, double *c);
void sqo_bg_energy_q_f64acc(int N0, int N1, int m, const double *h0, const double *h1,
                            const double *J, double c, const int8_t *q0, const int8_t *q1,
                            double *E);                               /* SharedFormulas.cpp:184-202 */

/* ---- dense brute-force search (CPUDenseGraphBFSearcher.cpp, CPUDenseGraphBatchSearch.cpp) ---- */
/* Single-worker semantics: ascending scan of x in [0, 2^N); strict < resets the
 * list, == appends while size < tileSize; result sorted ascending (:104-139).
 * Energies are evaluated exactly in double and compared after rounding to the
 * solver's real type (dtype 0 float, 1 double).  Emin starts at FLT_MAX.
 * Returns the number of solutions written to xs (packed, bit N-1-pos = x[pos]). */
int sqo_dg_bf_search(int dtype, int N, const double *W, int om, int tileSize,
                     double *Emin, unsigned long long *xs);
/* bipartite brute-force search (single-worker order); returns the number of
 * pairs (<= N0 + N1) written to xs0/xs1, -1 for unsupported sizes */
int sqo_bg_bf_search(int dtype, int N0, int N1, const double *b0, const double *b1, const double *W, int om,
                     int tileSize0, int tileSize1, double *EminOut, unsigned long long *xs0,
                     unsigned long long *xs1);

/* ---- expf ---- */
/* sum over float bit patterns b in [lo, hi) of splitmix64((b << 32) | bits(expf(b))),
 * host C library expf, NaN canonicalised to 0x7fc00000 */
unsigned long long sqo_expf_checksum(unsigned long long lo, unsigned long long hi);

#ifdef __cplusplus
}
#endif
#endif
