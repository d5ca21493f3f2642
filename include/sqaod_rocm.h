/*
 * sqaod-rocm C ABI: the rocm device of sqaod's simulated-quantum-annealing
 * solvers, for any FFI (the Python package binds it with ctypes).
 *
 * Each entry point replaces one function of the reference's per-device
 * CPython extension (the method tables of sqaodc/pyglue/annealer.inc:885-908,
 * bf_searcher.inc:479-494 and the device module sqaodpy/sqaod/cuda/src/
 * cuda_device.cpp), with the same name stem and argument meaning; numpy
 * arrays become plain pointers plus sizes, the dtype argument becomes the
 * `dtype` given at creation (0 = float32, 1 = float64).
 *
 * Conventions
 *  - Every function returns 0 on success and -1 on error; the message of the
 *    last error on the calling thread is returned by sqaod_rocm_last_error().
 *    Errors are the reference's RuntimeError cases (pyglue.h:394-399): wrong
 *    call order (state machine, Solver.cpp:96-112), shape mismatches, HIP
 *    failures.
 *  - Matrices are dense row-major without padding.  Spins q and bits x are
 *    int8 (q in {-1, +1}, x in {0, 1}).
 *  - Outputs are written into caller-provided buffers of the stated size.
 *  - Solvers are single-owner objects: calls on one handle must not overlap.
 */
#ifndef SQAOD_ROCM_H
#define SQAOD_ROCM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sqaod_rocm_device sqaod_rocm_device;
typedef struct sqaod_rocm_dg_annealer sqaod_rocm_dg_annealer;
typedef struct sqaod_rocm_dg_bf_searcher sqaod_rocm_dg_bf_searcher;
typedef struct sqaod_rocm_bg_annealer sqaod_rocm_bg_annealer;
typedef struct sqaod_rocm_bg_bf_searcher sqaod_rocm_bg_bf_searcher;
typedef struct sqaod_rocm_formulas sqaod_rocm_formulas;
typedef struct sqaod_rocm_dense_batch sqaod_rocm_dense_batch;
typedef struct sqaod_rocm_server sqaod_rocm_server;

/* ---- library ---- */

/* Message of the last failed call on this thread ("" if none). */
const char *sqaod_rocm_last_error(void);
/* Version probe; replaces sqaodc_cuda_version(int *version, int *cuda_version)
 * (sqaodc/cuda/api_cuda.cpp:5-9), found by name the way
 * sqaodpy/sqaod/common/envcheck.py:27-37 probes the CUDA library:
 * *version = the sqaodc API version (10003 = 1.0.3, the SQAODC_VERSION
 * encoding of sqaodc/common/defines.h:80), *hip_version = HIP_VERSION of the
 * HIP the library was built with (major * 10^7 + minor * 10^5 + patch). */
void sqaodc_rocm_version(int *version, int *hip_version);
/* Number of visible GPUs (sqaod.is_cuda_available counterpart, Common.cpp:20-77). */
int sqaod_rocm_device_count(int *count);

/* ---- device (cuda_device.cpp new/initialize/finalize/delete) ---- */
int sqaod_rocm_device_new(int devno, sqaod_rocm_device **out);
int sqaod_rocm_device_delete(sqaod_rocm_device *dev);
int sqaod_rocm_device_synchronize(sqaod_rocm_device *dev);

/* ---- dense graph annealer: annealer.inc method table (:885-908) ---- */

/* annealer_new (:14-30) */
int sqaod_rocm_dg_annealer_new(int dtype, sqaod_rocm_dg_annealer **out);
/* annealer_delete (:33-51) */
int sqaod_rocm_dg_annealer_delete(sqaod_rocm_dg_annealer *a);
/* annealer_assign_device (:54-74); without it the default device 0 is used */
int sqaod_rocm_dg_annealer_assign_device(sqaod_rocm_dg_annealer *a, sqaod_rocm_device *dev);
/* annealer_seed (:77-105) */
int sqaod_rocm_dg_annealer_seed(sqaod_rocm_dg_annealer *a, unsigned long long seed);
/* annealer_set_qubo (:108-141): W is N x N, symmetric; optimize 0 = minimize, 1 = maximize */
int sqaod_rocm_dg_annealer_set_qubo(sqaod_rocm_dg_annealer *a, const void *W, int N, int optimize);
/* annealer_set_hamiltonian (:144-160): h[N], J[N x N] symmetric, c scalar (all of dtype) */
int sqaod_rocm_dg_annealer_set_hamiltonian(sqaod_rocm_dg_annealer *a, const void *h, const void *J, const void *c,
                                           int N);
/* annealer_get_problem_size (:163-196) */
int sqaod_rocm_dg_annealer_get_problem_size(sqaod_rocm_dg_annealer *a, int *N);
/* annealer_set_preferences (:279-299), one preference per call, value as text:
 * "algorithm" ("coloring" | "naive" | "sa_naive" | "default" | "sa_default";
 * others fall back to the default of their kind, Solver.cpp:152-168),
 * "n_trotters" (int), and the rocm-device knobs "rng" ("philox" | "mt19937"),
 * "n_workers" (int >= 1: with rng = mt19937, the per-worker streams of the
 * CPU annealer's OpenMP coloring step with that many threads),
 * "sweep" ("local" | "plain"), "field_precision" ("fp32" | "bf16x2" | "fp8"),
 * "multi_flip" ("auto" | "on" | "off" | a number 2..64: float local-field
 * sweeps commit every flip of a 64-attempt batch in order, instead of ending
 * the batch at its first flip, after a batch with at least 4 (or the given
 * number of) accepting lanes; "auto" does so in the first 3 sweeps after the
 * spins were set, when most attempts are accepted; the results are the same
 * in every mode),
 * "helper_wave" ("on" | "off": single-flip float local-field sweeps with one
 * slice per workgroup run a second wave per slice that applies the deferred
 * J rows to the fields; the results are the same either way),
 * "slices_per_block" (int, 0 = auto), "sweep_launch" ("auto" | "cooperative"),
 * "refresh_interval" (int, steps between
 * field recomputes; 0 = only after the spins were set), "refresh_engine"
 * ("mfma" | "vector"), "profile" ("0" | "1"), "spin_limit" (int > 0: blocked
 * iterations a slice may wait on a neighbour's decision before the sweep gives
 * up; the step then fails with an error instead of hanging, and the spins must
 * be set again). */
int sqaod_rocm_dg_annealer_set_preference(sqaod_rocm_dg_annealer *a, const char *name, const char *value);
/* annealer_get_preferences (:302-329): "name=value" pairs joined by ';'.
 * Writes at most buflen bytes (NUL-terminated); *needed receives the full length + 1. */
int sqaod_rocm_dg_annealer_get_preferences(sqaod_rocm_dg_annealer *a, char *buf, size_t buflen, size_t *needed);
/* number of Trotter slices m (the length of get_E / rows of get_x, get_q) */
int sqaod_rocm_dg_annealer_get_n_trotters(sqaod_rocm_dg_annealer *a, int *m);
/* annealer_get_E (:332-364): E[m] of dtype */
int sqaod_rocm_dg_annealer_get_E(sqaod_rocm_dg_annealer *a, void *E, int m);
/* annealer_get_x (:367-392): x[m x N] int8 */
int sqaod_rocm_dg_annealer_get_x(sqaod_rocm_dg_annealer *a, int8_t *x, int m, int N);
/* annealer_get_hamiltonian (:470-503): h[N], J[N x N], c[1] of dtype */
int sqaod_rocm_dg_annealer_get_hamiltonian(sqaod_rocm_dg_annealer *a, void *h, void *J, void *c);
/* annealer_get_q (:506-548): q[m x N] int8 */
int sqaod_rocm_dg_annealer_get_q(sqaod_rocm_dg_annealer *a, int8_t *q, int m, int N);
/* annealer_set_q (:395-431): one q[N] for every slice */
int sqaod_rocm_dg_annealer_set_q(sqaod_rocm_dg_annealer *a, const int8_t *q, int N);
/* annealer_set_qset (:434-467): q[m x N], sets m */
int sqaod_rocm_dg_annealer_set_qset(sqaod_rocm_dg_annealer *a, const int8_t *q, int m, int N);
/* annealer_radomize_q (:747-763) */
int sqaod_rocm_dg_annealer_randomize_spin(sqaod_rocm_dg_annealer *a);
/* annealer_calculate_E (:766-783) */
int sqaod_rocm_dg_annealer_calculate_E(sqaod_rocm_dg_annealer *a);
/* annealer_prepare (:786-802) */
int sqaod_rocm_dg_annealer_prepare(sqaod_rocm_dg_annealer *a);
/* annealer_make_solution (:805-833) */
int sqaod_rocm_dg_annealer_make_solution(sqaod_rocm_dg_annealer *a);
/* annealer_get_system_E (:836-850): G, beta converted to dtype; result as double */
int sqaod_rocm_dg_annealer_get_system_E(sqaod_rocm_dg_annealer *a, double G, double beta, double *E);
/* annealer_anneal_one_step (:853-878): G, beta converted to dtype */
int sqaod_rocm_dg_annealer_anneal_one_step(sqaod_rocm_dg_annealer *a, double G, double beta);
/* Diagnostics without a reference counterpart: cumulative accepted flips,
 * committed batches and blocked waits of the sweeps since prepare(), and the
 * max |E_dtype - E_fp64| of the last energy evaluation. */
int sqaod_rocm_dg_annealer_get_counters(sqaod_rocm_dg_annealer *a, unsigned long long counters[3]);
int sqaod_rocm_dg_annealer_get_energy_rounding_error(sqaod_rocm_dg_annealer *a, double *err);
/* Diagnostic without a reference counterpart: E[m] of the last energy
 * evaluation (calculate_E / make_solution) in double, before its rounding to
 * dtype.  Exact on integer problems; otherwise within FP64 rounding of the
 * exact value for the J held in dtype. */
int sqaod_rocm_dg_annealer_get_E_fp64(sqaod_rocm_dg_annealer *a, double *E, int m);
/* System-energy trace, for callers that record getSystemE after every step
 * (sqaodpy/sqaod/wildqat/opt.py:33-35, 57-62 calls get_system_E per step);
 * no reference counterpart.  With capacity > 0, every anneal_one_step(G,
 * beta) appends the value get_system_E(G, beta) would return for the state
 * after the step, computed on the device without a host synchronisation
 * (the first `capacity` steps are kept); 0 turns the trace off.  Setting it
 * clears the trace.  get_system_E_trace copies up to cap kept values (as
 * double) into out, *n = how many, *steps = steps traced since it was set. */
int sqaod_rocm_dg_annealer_set_system_E_trace(sqaod_rocm_dg_annealer *a, int capacity);
int sqaod_rocm_dg_annealer_get_system_E_trace(sqaod_rocm_dg_annealer *a, double *out, int cap, int *n,
                                              long long *steps);
/* Diagnostic without a reference counterpart: max over slices of
 * |E_form(q_y) - E(q_y)|, both in FP64, for the current spins, where E_form
 * uses the J the field precision form anneals (bf16 hi + lo or the fp8 J')
 * and E the J as given; 0 for the fp32/fp64 form. */
int sqaod_rocm_dg_annealer_get_form_energy_error(sqaod_rocm_dg_annealer *a, double *err);
/* Diagnostic without a reference counterpart: the local fields F[y][i] =
 * sum_j J[i][j] q[y][j] (m x N, dtype) the local-field sweep holds, to check
 * the field GEMM and its precision forms. */
int sqaod_rocm_dg_annealer_get_fields(sqaod_rocm_dg_annealer *a, void *F, int m, int N);
/* With preference "profile" = "1": device time (ms, HIP events on the solver's
 * stream) and launch counts of [field refresh, random-stream generation, sweep]
 * accumulated since the previous call; resets the accumulators. */
int sqaod_rocm_dg_annealer_get_phase_times(sqaod_rocm_dg_annealer *a, double ms[3], unsigned long long counts[3]);

/* ---- dense graph brute-force searcher: bf_searcher.inc method table (:479-494) ----
 * Result: Emin = min over x in [0, 2^N) of x^T W x (maximize: max), and the
 * first tile_size x reaching it in ascending order of the packed x (bit N-1-i
 * of the packed x is x[i], Common.cpp:79-93) -- the sqaod.cpu one-worker
 * result (CPUDenseGraphBatchSearch.cpp:26-50, CPUDenseGraphBFSearcher.cpp:
 * 104-139), which the CUDA device's tile merge also gives. */
/* bf_searcher_new / delete / assign_device (:479-482) */
int sqaod_rocm_dg_bf_searcher_new(int dtype, sqaod_rocm_dg_bf_searcher **out);
int sqaod_rocm_dg_bf_searcher_delete(sqaod_rocm_dg_bf_searcher *b);
int sqaod_rocm_dg_bf_searcher_assign_device(sqaod_rocm_dg_bf_searcher *b, sqaod_rocm_device *dev);
/* bf_searcher_set_qubo (:75-104): W N x N symmetric, N <= 63; optimize 0 = minimize, 1 = maximize */
int sqaod_rocm_dg_bf_searcher_set_qubo(sqaod_rocm_dg_bf_searcher *b, const void *W, int N, int optimize);
/* bf_searcher_get_problem_size (:107-128) */
int sqaod_rocm_dg_bf_searcher_get_problem_size(sqaod_rocm_dg_bf_searcher *b, int *N);
/* bf_searcher_set_preferences (:486): "tile_size" (rounded up to a multiple of
 * 256, Solver.cpp:201-209; caps the solution list), "algorithm" (always
 * brute_force_search) */
int sqaod_rocm_dg_bf_searcher_set_preference(sqaod_rocm_dg_bf_searcher *b, const char *name, const char *value);
/* bf_searcher_get_preferences (:487), format as for the annealer */
int sqaod_rocm_dg_bf_searcher_get_preferences(sqaod_rocm_dg_bf_searcher *b, char *buf, size_t buflen,
                                              size_t *needed);
/* bf_searcher_prepare / calculate_E / make_solution (:490-492) */
int sqaod_rocm_dg_bf_searcher_prepare(sqaod_rocm_dg_bf_searcher *b);
int sqaod_rocm_dg_bf_searcher_calculate_E(sqaod_rocm_dg_bf_searcher *b);
int sqaod_rocm_dg_bf_searcher_make_solution(sqaod_rocm_dg_bf_searcher *b);
/* bf_searcher_search_range (:424-445): scans the next span of x on the device
 * (max(tile_size, 2^26) candidates; the reference's spans are one tile, or one
 * tile per CPU worker); *done = 1 once all 2^N are scanned, *cur_x_end = end */
int sqaod_rocm_dg_bf_searcher_search_range(sqaod_rocm_dg_bf_searcher *b, int *done, unsigned long long *cur_x_end);
/* bf_searcher_search (:456-473): prepare, scan everything, make_solution */
int sqaod_rocm_dg_bf_searcher_search(sqaod_rocm_dg_bf_searcher *b);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the search of sqaod_rocm_dg_bf_searcher_search (same
 * kernels, same minimum and x list) on the caller's device buffers, with
 * every intermediate in a static 256 KiB device array of the library (one
 * per GPU, held until the process ends); no allocation or free.
 *   d_W: N x N symmetric QUBO of dtype (0 float32, 1 float64) in device
 *   memory, N in 1..63; optimize 0 = minimize, 1 = maximize;
 *   tile_size in 1..2^26, rounded up to a multiple of 256 and capped at 2^N
 *   as the searcher does -- d_x must hold that many entries;
 *   d_Emin: one value of dtype; d_x: the packed minimisers in ascending order
 *   (bit N-1-i = x[i]); d_n: their number (unsigned long long);
 *   stream: the hipStream_t to enqueue on (NULL: the null stream).
 * Asynchronous: the results are ready when the stream reaches them.  Calls
 * that overlap on one GPU share the scratch array and race. */
int sqaod_rocm_dense_bf_search_noalloc(int dtype, const void *d_W, int N, int optimize, unsigned long long tile_size,
                                       void *d_Emin, unsigned long long *d_x, unsigned long long *d_n, void *stream);
/* number of solutions (rows of get_x); get_E has max(1, n) entries */
int sqaod_rocm_dg_bf_searcher_get_n_solutions(sqaod_rocm_dg_bf_searcher *b, int *n);
/* bf_searcher_get_x (:241-275): x[n x N] int8 */
int sqaod_rocm_dg_bf_searcher_get_x(sqaod_rocm_dg_bf_searcher *b, int8_t *x, int n, int N);
/* bf_searcher_get_E (:318-341): E[n] of dtype, n = max(1, number of solutions) */
int sqaod_rocm_dg_bf_searcher_get_E(sqaod_rocm_dg_bf_searcher *b, void *E, int n);

/* ---- bipartite graph brute-force searcher: bf_searcher.inc method table
 * (:479-494), the BIPARTITE_GRAPH variants.  W is N1 x N0.  Result: Emin = min
 * over all (x0, x1) of b0.x0 + b1.x1 + x1^T W x0 (maximize: max), and the
 * first N0 + N1 pairs reaching it in one worker's visiting order (x0 tile of
 * tile_size_0, then x1, then x0 within the tile;
 * CPUBipartiteGraphBatchSearch.cpp:29-62, CPUBipartiteGraphBFSearcher.cpp:
 * 166-198). ---- */
int sqaod_rocm_bg_bf_searcher_new(int dtype, sqaod_rocm_bg_bf_searcher **out);
int sqaod_rocm_bg_bf_searcher_delete(sqaod_rocm_bg_bf_searcher *b);
int sqaod_rocm_bg_bf_searcher_assign_device(sqaod_rocm_bg_bf_searcher *b, sqaod_rocm_device *dev);
/* bf_searcher_set_qubo / get_problem_size (:133-192): b0[N0], b1[N1], W[N1 x N0], N0, N1 <= 63 */
int sqaod_rocm_bg_bf_searcher_set_qubo(sqaod_rocm_bg_bf_searcher *b, const void *b0, int N0, const void *b1, int N1,
                                       const void *W, int optimize);
int sqaod_rocm_bg_bf_searcher_get_problem_size(sqaod_rocm_bg_bf_searcher *b, int *N0, int *N1);
/* "tile_size_0", "tile_size_1" (rounded up to multiples of 256, Solver.cpp:231-251) */
int sqaod_rocm_bg_bf_searcher_set_preference(sqaod_rocm_bg_bf_searcher *b, const char *name, const char *value);
int sqaod_rocm_bg_bf_searcher_get_preferences(sqaod_rocm_bg_bf_searcher *b, char *buf, size_t buflen,
                                              size_t *needed);
int sqaod_rocm_bg_bf_searcher_prepare(sqaod_rocm_bg_bf_searcher *b);
int sqaod_rocm_bg_bf_searcher_calculate_E(sqaod_rocm_bg_bf_searcher *b);
int sqaod_rocm_bg_bf_searcher_make_solution(sqaod_rocm_bg_bf_searcher *b);
/* bf_searcher_search_range (:429-451): scans the next rectangle of pairs;
 * *done = 1 once all are scanned; *cur_x0, *cur_x1: the position reached */
int sqaod_rocm_bg_bf_searcher_search_range(sqaod_rocm_bg_bf_searcher *b, int *done, unsigned long long *cur_x0,
                                           unsigned long long *cur_x1);
int sqaod_rocm_bg_bf_searcher_search(sqaod_rocm_bg_bf_searcher *b);
int sqaod_rocm_bg_bf_searcher_get_n_solutions(sqaod_rocm_bg_bf_searcher *b, int *n);
/* bf_searcher_get_x (:277-319): x0[n x N0], x1[n x N1] int8 */
int sqaod_rocm_bg_bf_searcher_get_x(sqaod_rocm_bg_bf_searcher *b, int8_t *x0, int8_t *x1, int n, int N0, int N1);
/* bf_searcher_get_E (:321-346): E[max(1, n)] of dtype */
int sqaod_rocm_bg_bf_searcher_get_E(sqaod_rocm_bg_bf_searcher *b, void *E, int n);

/* ---- bipartite graph annealer: annealer.inc method table (:885-908), the
 * BIPARTITE_GRAPH variants.  W is N1 x N0 (rows index the x1 side), J likewise;
 * q/x pairs are passed as two arrays (m x N0 and m x N1 int8). ---- */
/* annealer_new / delete / assign_device / seed (:14-105) */
int sqaod_rocm_bg_annealer_new(int dtype, sqaod_rocm_bg_annealer **out);
int sqaod_rocm_bg_annealer_delete(sqaod_rocm_bg_annealer *a);
int sqaod_rocm_bg_annealer_assign_device(sqaod_rocm_bg_annealer *a, sqaod_rocm_device *dev);
int sqaod_rocm_bg_annealer_seed(sqaod_rocm_bg_annealer *a, unsigned long long seed);
/* annealer_set_qubo (:108-141, bipartite form): b0[N0], b1[N1], W[N1 x N0] */
int sqaod_rocm_bg_annealer_set_qubo(sqaod_rocm_bg_annealer *a, const void *b0, int N0, const void *b1, int N1,
                                    const void *W, int optimize);
/* annealer_set_hamiltonian (:144-160): h0[N0], h1[N1], J[N1 x N0], c scalar */
int sqaod_rocm_bg_annealer_set_hamiltonian(sqaod_rocm_bg_annealer *a, const void *h0, int N0, const void *h1, int N1,
                                           const void *J, const void *c);
/* annealer_get_problem_size (:163-196): (N0, N1) */
int sqaod_rocm_bg_annealer_get_problem_size(sqaod_rocm_bg_annealer *a, int *N0, int *N1);
/* annealer_set_preferences / get_preferences (:279-329): "algorithm"
 * ("coloring" | "sa_coloring", other names fall back as on the CUDA device),
 * "n_trotters", "rng" ("philox" | "mt19937"), "field_precision" ("fp32" |
 * "bf16x2"), "refresh_engine" ("mfma" | "vector") */
int sqaod_rocm_bg_annealer_set_preference(sqaod_rocm_bg_annealer *a, const char *name, const char *value);
int sqaod_rocm_bg_annealer_get_preferences(sqaod_rocm_bg_annealer *a, char *buf, size_t buflen, size_t *needed);
/* number of Trotter slices m */
int sqaod_rocm_bg_annealer_get_n_trotters(sqaod_rocm_bg_annealer *a, int *m);
/* annealer_get_E (:332-364): E[m] of dtype */
int sqaod_rocm_bg_annealer_get_E(sqaod_rocm_bg_annealer *a, void *E, int m);
/* Diagnostic without a reference counterpart: E[m] of the last energy
 * evaluation in double, before its rounding to dtype. */
int sqaod_rocm_bg_annealer_get_E_fp64(sqaod_rocm_bg_annealer *a, double *E, int m);
/* annealer_get_x (:367-392): x0[m x N0], x1[m x N1] int8 */
int sqaod_rocm_bg_annealer_get_x(sqaod_rocm_bg_annealer *a, int8_t *x0, int8_t *x1, int m, int N0, int N1);
/* annealer_get_hamiltonian (:470-503): h0[N0], h1[N1], J[N1 x N0], c[1] of dtype */
int sqaod_rocm_bg_annealer_get_hamiltonian(sqaod_rocm_bg_annealer *a, void *h0, void *h1, void *J, void *c);
/* annealer_get_q (:506-548): q0[m x N0], q1[m x N1] int8 */
int sqaod_rocm_bg_annealer_get_q(sqaod_rocm_bg_annealer *a, int8_t *q0, int8_t *q1, int m, int N0, int N1);
/* annealer_set_q (:395-431, (q0, q1) tuple): one pair for every slice */
int sqaod_rocm_bg_annealer_set_q(sqaod_rocm_bg_annealer *a, const int8_t *q0, int N0, const int8_t *q1, int N1);
/* annealer_set_qset (:434-467): m pairs, sets m */
int sqaod_rocm_bg_annealer_set_qset(sqaod_rocm_bg_annealer *a, const int8_t *q0, const int8_t *q1, int m, int N0,
                                    int N1);
/* annealer_radomize_q / calculate_E / prepare / make_solution (:747-833) */
int sqaod_rocm_bg_annealer_randomize_spin(sqaod_rocm_bg_annealer *a);
int sqaod_rocm_bg_annealer_calculate_E(sqaod_rocm_bg_annealer *a);
int sqaod_rocm_bg_annealer_prepare(sqaod_rocm_bg_annealer *a);
int sqaod_rocm_bg_annealer_make_solution(sqaod_rocm_bg_annealer *a);
/* annealer_get_system_E (:836-850), annealer_anneal_one_step (:853-878) */
int sqaod_rocm_bg_annealer_get_system_E(sqaod_rocm_bg_annealer *a, double G, double beta, double *E);
int sqaod_rocm_bg_annealer_anneal_one_step(sqaod_rocm_bg_annealer *a, double G, double beta);

/* ---- formulas: formulas.inc method table (:586-606).  One object serves the
 * dense and the bipartite functions (dg_formulas_new and bg_formulas_new both
 * map to sqaod_rocm_formulas_new); dtype per call as in the reference.  The
 * single-vector functions are the batch functions with nBatch = 1.  Energies
 * are formed on the device in FP64 and rounded to dtype once; Hamiltonians on
 * the host in double. ---- */
/* dg_formulas_new / bg_formulas_new, _delete, _assign_device (:27-73, :280-330) */
int sqaod_rocm_formulas_new(sqaod_rocm_formulas **out);
int sqaod_rocm_formulas_delete(sqaod_rocm_formulas *f);
int sqaod_rocm_formulas_assign_device(sqaod_rocm_formulas *f, sqaod_rocm_device *dev);
/* dense_graph_(batch_)calculate_E (:89-157): E[b] = x_b^T W x_b, x int8 nBatch x N */
int sqaod_rocm_dense_graph_batch_calculate_E(sqaod_rocm_formulas *f, int dtype, void *E, const void *W, int N,
                                             const int8_t *x, int nBatch);
/* dense_graph_calculate_hamiltonian (:159-201): h[N], J[N x N], c[1] */
int sqaod_rocm_dense_graph_calculate_hamiltonian(sqaod_rocm_formulas *f, int dtype, void *h, void *J, void *c,
                                                 const void *W, int N);
/* dense_graph_(batch_)calculate_E_from_spin (:203-276): E[b] = -c - h.q_b - q_b^T J q_b */
int sqaod_rocm_dense_graph_batch_calculate_E_from_spin(sqaod_rocm_formulas *f, int dtype, void *E, const void *h,
                                                       const void *J, const void *c, const int8_t *q, int N,
                                                       int nBatch);
/* bipartite_graph_(batch_)calculate_E (:346-420): E[b] = b0.x0_b + b1.x1_b + x1_b^T W x0_b, W N1 x N0 */
int sqaod_rocm_bipartite_graph_batch_calculate_E(sqaod_rocm_formulas *f, int dtype, void *E, const void *b0, int N0,
                                                 const void *b1, int N1, const void *W, const int8_t *x0,
                                                 const int8_t *x1, int nBatch);
/* bipartite_graph_batch_calculate_E_2d (:422-462): E[nBatch1 x nBatch0] over all (x0, x1) pairs */
int sqaod_rocm_bipartite_graph_batch_calculate_E_2d(sqaod_rocm_formulas *f, int dtype, void *E, const void *b0,
                                                    int N0, const void *b1, int N1, const void *W, const int8_t *x0,
                                                    int nBatch0, const int8_t *x1, int nBatch1);
/* bipartite_graph_calculate_hamiltonian (:464-510): h0[N0], h1[N1], J[N1 x N0], c[1] */
int sqaod_rocm_bipartite_graph_calculate_hamiltonian(sqaod_rocm_formulas *f, int dtype, void *h0, void *h1, void *J,
                                                     void *c, const void *b0, int N0, const void *b1, int N1,
                                                     const void *W);
/* bipartite_graph_(batch_)calculate_E_from_spin (:512-578): E[b] = -(h0.q0 + h1.q1 + q1^T J q0) - c */
int sqaod_rocm_bipartite_graph_batch_calculate_E_from_spin(sqaod_rocm_formulas *f, int dtype, void *E, const void *h0,
                                                           int N0, const void *h1, int N1, const void *J,
                                                           const void *c, const int8_t *q0, const int8_t *q1,
                                                           int nBatch);

/* Self check without a reference counterpart: order-independent checksum of
 * the device exp (the acceptance test's expf) over float bit patterns
 * [lo, hi), for comparison with the host C library's expf:
 * sum of splitmix64((b << 32) | bits(expf(b))), NaN canonicalised. */
int sqaod_rocm_expf_checksum(unsigned long long lo, unsigned long long hi, unsigned long long *out);

/* ---------------- dense batch (no counterpart in the reference) ----------------
 * For callers that anneal many independent dense QUBOs at once (hundreds of
 * instances of uneven N and m): every instance is a float32 dense annealer
 * (Philox streams, local-field sweep, MFMA refresh), and a step of the batch
 * runs all of them in shared launches -- one grouped field GEMM, one grouped
 * stream generation and one persistent sweep over every (instance, slice) --
 * optionally captured as one hipGraph for a whole schedule.  Each instance's
 * spins and energies are identical to a sqaod_rocm_dg_annealer run alone with
 * the same seed, preferences and schedule. */
int sqaod_rocm_dense_batch_new(sqaod_rocm_dense_batch **out);
int sqaod_rocm_dense_batch_delete(sqaod_rocm_dense_batch *b);
/* field_precision | refresh_interval | spin_limit | algorithm (applied to every
 * instance), waves_per_block (0 = tuned, 1, 2, 4), order (largest_first | given) */
int sqaod_rocm_dense_batch_set_preference(sqaod_rocm_dense_batch *b, const char *name, const char *value);
/* W: N x N float32 symmetric QUBO; maximize != 0 maximizes; *index: the instance */
int sqaod_rocm_dense_batch_add_qubo(sqaod_rocm_dense_batch *b, const float *W, int N, int m, unsigned long long seed,
                                    int maximize, int *index);
int sqaod_rocm_dense_batch_size(sqaod_rocm_dense_batch *b, int *n);
/* prepare + randomize_spin of every instance */
int sqaod_rocm_dense_batch_prepare(sqaod_rocm_dense_batch *b);
/* nSteps steps with (G[k], beta[k]); graph != 0: one hipGraph for the whole
 * schedule; *ms (may be null): device time of the schedule */
int sqaod_rocm_dense_batch_anneal(sqaod_rocm_dense_batch *b, const double *G, const double *beta, int nSteps,
                                  int graph, double *ms);
/* times trialSteps steps at (G, beta) per candidate launch shape and keeps the
 * fastest; spins, fields and streams are restored afterwards */
int sqaod_rocm_dense_batch_autotune(sqaod_rocm_dense_batch *b, double G, double beta, int trialSteps,
                                    int *wavesPerBlock, int *largestFirst);
int sqaod_rocm_dense_batch_attempts_per_step(sqaod_rocm_dense_batch *b, long long *n);
/* instance i: its shape, FP64-recomputed energies (m), solutions x (m x N) and spins q (m x N) */
int sqaod_rocm_dense_batch_get_shape(sqaod_rocm_dense_batch *b, int i, int *N, int *m);
int sqaod_rocm_dense_batch_get_E(sqaod_rocm_dense_batch *b, int i, float *E, int m);
int sqaod_rocm_dense_batch_get_x(sqaod_rocm_dense_batch *b, int i, int8_t *x, int m, int N);
int sqaod_rocm_dense_batch_get_q(sqaod_rocm_dense_batch *b, int i, int8_t *q, int m, int N);

/* ---------------- serving (no counterpart in the reference) ----------------
 * For callers that serve anneal requests as they arrive: jobs (dense or
 * bipartite float32 QUBOs, each with its own seed and G/beta schedule) are
 * admitted in arrival order while an active-job limit and a device-memory
 * budget allow; every tick advances all active jobs by one step together
 * (dense jobs in shared launches, bipartite jobs on parallel streams) and
 * finishes the jobs whose schedule is done.  A job's E and x equal the same
 * job run alone with the same seed and schedule. */
int sqaod_rocm_server_new(sqaod_rocm_server **out);
int sqaod_rocm_server_delete(sqaod_rocm_server *s);
/* memory_budget_mb, max_active, dense_field_precision (fp32 | bf16x2 | fp8),
 * bipartite_field_precision (fp32 | bf16x2), bipartite_streams, block_fraction,
 * dense_spin_limit (the bounded wait of dense jobs submitted afterwards; 0 = default) */
int sqaod_rocm_server_set_preference(sqaod_rocm_server *s, const char *name, const char *value);
/* W: N x N symmetric (dense) or N1 x N0 (bipartite, b0[N0], b1[N1]); minimized.
 * A job over the memory budget, or a dense job whose N or m the grouped sweep
 * cannot hold (LDS region per slice, resident waves), is rejected here. */
int sqaod_rocm_server_submit_dense(sqaod_rocm_server *s, const float *W, int N, int m, unsigned long long seed,
                                   const double *G, const double *beta, int nSteps, int *id);
int sqaod_rocm_server_submit_bipartite(sqaod_rocm_server *s, const float *b0, const float *b1, const float *W, int N0,
                                       int N1, int m, unsigned long long seed, const double *G, const double *beta,
                                       int nSteps, int *id);
/* one step of every active job; *remaining = active + pending jobs after it */
int sqaod_rocm_server_tick(sqaod_rocm_server *s, int *remaining);
/* pending, active, bytes in use and the peak, by the admission estimate */
int sqaod_rocm_server_status(sqaod_rocm_server *s, int *pending, int *active, unsigned long long *bytesInUse,
                             unsigned long long *peakBytes);
/* ids of the jobs finished since the last call (up to cap; *n written) */
int sqaod_rocm_server_take_done(sqaod_rocm_server *s, int *ids, int cap, int *n);
int sqaod_rocm_server_result_shape(sqaod_rocm_server *s, int id, int *kind, int *N0, int *N1, int *m);
/* E[m]; x0[m x N0] (dense: x), x1[m x N1] (bipartite, else may be null);
 * times: submit, admit, finish (seconds on the server clock); attempts of the job.
 * Fails (-1, the job's error as the message) for a failed job. */
int sqaod_rocm_server_result(sqaod_rocm_server *s, int id, float *E, int8_t *x0, int8_t *x1, double times[3],
                             long long *attempts);
/* *failed = 1 when the job failed -- its admission, a step or its solution
 * threw (a HIP error, an aborted sweep) -- with the message in buf (at most
 * buflen bytes, NUL-terminated); the server goes on with the other jobs. */
int sqaod_rocm_server_result_error(sqaod_rocm_server *s, int id, char *buf, size_t buflen, int *failed);
int sqaod_rocm_server_release(sqaod_rocm_server *s, int id);
int sqaod_rocm_server_now(sqaod_rocm_server *s, double *t);

/* Self check without a reference counterpart: the fp8 sweep's e4m3 decode
 * (gfx950 v_cvt_pk_f32_fp8) applied to the byte codes 0..255 (out[256]). */
int sqaod_rocm_fp8_decode_table(float *out);

#ifdef __cplusplus
}
#endif
#endif
