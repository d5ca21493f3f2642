"""The reference's dense-annealer test intents (sqaodpy/tests/
test_dense_graph_annealer.py) for the rocm device, in float32 and float64:
problem size, preferences, the Hamiltonian round trip (:57-82), energies of
set spins (:84-120), set_q / set_qset (:122-152), get_E / get_x of ones
(:25-50, :154-165), annealing trivial problems to 0 with naive, coloring,
default and SA (:167-265), algorithm selection (:267-290, :404-420), the
system energy identities (:292-354), solver reuse (:356-361) and the error
before prepare (:388-393); then the device's system-energy trace against
per-step get_system_E and the oracle, and the wildqat adapter
(sqaodpy/sqaod/wildqat/opt.py) on top of it."""
from math import exp, log

import numpy as np
import pytest

from problems import SEED_BASE, dense_graph_random, dense_int_W, schedule

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64]


def epu(dtype):
    return 1.e-6 if dtype == np.float32 else 1.e-12


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


@pytest.fixture(scope="module")
def sq():
    import sqaod_amd as sq
    return sq


def new_annealer(rocm, N, m, dtype, seed=SEED_BASE):
    an = rocm.dense_graph_annealer(dtype=dtype)
    an.set_qubo(dense_graph_random(N, seed, dtype))
    an.set_preferences(n_trotters=m)
    return an


def hamiltonian(W):
    """SharedFormulas.cpp:36-55 in float64: h = -1/2 colsum W, J = -1/4 W, c
    = sum J + trace J, then the diagonal of J zeroed."""
    W = np.asarray(W, np.float64)
    h = -0.5 * W.sum(axis=0)
    J = -0.25 * W
    c = J.sum() + np.trace(J)
    J = J.copy()
    np.fill_diagonal(J, 0.)
    return h, J, c


@pytest.mark.parametrize("dtype", DTYPES)
def test_calling_sequence_problem_size_and_prefs(rocm, sq, dtype):
    N, m = 200, 100
    an = new_annealer(rocm, N, m, dtype)
    an.set_preferences(algorithm=sq.algorithm.default)
    an.seed(0)
    an.prepare()
    an.randomize_spin()
    an.anneal_one_step(1, 1)
    an.calculate_E()
    an.get_system_E(1, 1)
    an.make_solution()
    assert len(an.get_E()) == m
    assert an.get_problem_size() == N
    x = an.get_x()
    assert len(x) == m and x[0].shape == (N,)
    p = an.get_preferences()
    assert p["n_trotters"] == m and p["device"] == "rocm"
    assert p["precision"] == ("float" if dtype == np.float32 else "double")
    an.set_preferences(n_trotters=3)
    assert an.get_preferences()["n_trotters"] == 3
    assert an.dtype == dtype


@pytest.mark.parametrize("dtype", DTYPES)
def test_hamiltonian_round_trip(rocm, dtype):
    N = 100
    W = dense_graph_random(N, SEED_BASE + 1, dtype)
    an = rocm.dense_graph_annealer(dtype=dtype)
    an.set_qubo(W)
    h0, J0, c0 = an.get_hamiltonian()
    h1, J1, c1 = hamiltonian(W)
    assert np.allclose(h0, h1, atol=epu(dtype)) and np.allclose(J0, J1, atol=epu(dtype))
    assert np.allclose(c0, c1, atol=epu(dtype) * N * N)
    an2 = rocm.dense_graph_annealer(dtype=dtype)
    an2.set_hamiltonian(h1, J1, c1)
    an2.set_preferences(n_trotters=10)
    an2.prepare()
    h2, J2, c2 = an2.get_hamiltonian()
    assert np.array_equal(h2, h1.astype(dtype)) and np.array_equal(J2, J1.astype(dtype))
    assert c2 == dtype(c1)


@pytest.mark.parametrize("dtype", DTYPES)
def test_energies_of_set_spins(rocm, sq, dtype):
    N = 200
    an = new_annealer(rocm, N, 1, dtype)
    an.prepare()
    an.set_q(-np.ones(N, np.int8))
    assert np.allclose(an.get_E()[0], 0., atol=epu(dtype))
    N = 8
    W = dense_graph_random(N, SEED_BASE + 2, dtype)
    an = rocm.dense_graph_annealer(W, dtype=dtype, n_trotters=1)
    an.prepare()
    for i in range(1 << N):
        x = sq.create_bitset_sequence((i,), N)[0]
        Eref = float(x.astype(np.float64) @ W.astype(np.float64) @ x.astype(np.float64))
        an.set_q(sq.bit_to_spin(x))
        assert np.allclose(an.get_E()[0], Eref, atol=epu(dtype) * 10), i
    W = np.ones((N, N), dtype)
    an = rocm.dense_graph_annealer(W, dtype=dtype, n_trotters=1)
    an.prepare()
    an.set_q(np.ones(N, np.int8))
    assert an.get_E()[0] == N * N
    x = an.get_x()
    assert len(x) == 1 and np.all(x[0] == 1)


@pytest.mark.parametrize("dtype", DTYPES)
def test_set_q_and_set_qset(rocm, dtype):
    rng = np.random.default_rng(5)
    N = m = 100
    an = new_annealer(rocm, N, m, dtype)
    an.prepare()
    qin = (2 * rng.integers(0, 2, N) - 1).astype(np.int8)
    an.set_q(qin)
    qout = an.get_q()
    assert len(qout) == m and all(np.array_equal(qin, q) for q in qout)
    N, m = 5, 2
    an = new_annealer(rocm, N, m, dtype)
    qset = [(2 * rng.integers(0, 2, N) - 1).astype(np.int8) for _ in range(m)]
    an.set_qset(qset)
    qout = an.get_q()
    assert len(qout) == m and all(np.array_equal(a, b) for a, b in zip(qset, qout))


def anneal(an, sq):
    """test_dense_graph_annealer.py:52-73"""
    an.prepare()
    an.randomize_spin()
    algo = an.get_preferences()["algorithm"]
    if algo in (sq.algorithm.sa_naive, sq.algorithm.sa_coloring):
        Ginit, Gfin, beta = 10., 0.02, 1.
    else:
        Ginit, Gfin, beta = 5., 0.02, 1. / 0.03
    nSteps = 100
    G, tau = Ginit, exp(log(Gfin / Ginit) / nSteps)
    for _ in range(nSteps):
        an.anneal_one_step(G, beta)
        G *= tau


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("algo,m", [("naive", 4), ("coloring", 4), ("default", 4), ("default", 1),
                                    ("sa_naive", 1)])
def test_anneal_trivial_problems(rocm, sq, dtype, algo, m):
    N = 10
    an = rocm.dense_graph_annealer(np.ones((N, N), dtype), sq.minimize, dtype=dtype, n_trotters=m,
                                   algorithm=algo)
    an.seed(3)
    anneal(an, sq)
    assert an.get_E().min() == 0
    an = rocm.dense_graph_annealer(-np.ones((N, N), dtype), sq.maximize, dtype=dtype, n_trotters=m,
                                   algorithm=algo)
    an.seed(4)
    anneal(an, sq)
    assert an.get_E().max() == 0
    h, J, c = hamiltonian(np.ones((N, N)))
    an = rocm.dense_graph_annealer(dtype=dtype)
    an.set_hamiltonian(h, J, c)
    an.set_preferences(n_trotters=m, algorithm=algo)
    an.seed(5)
    anneal(an, sq)
    assert an.get_E().max() == 0
    anneal(an, sq)  # reuse: test_reuse_solver


@pytest.mark.parametrize("dtype", DTYPES)
def test_algorithm_selection(rocm, sq, dtype):
    an = new_annealer(rocm, 10, 10, dtype)
    for algo in (sq.algorithm.naive, sq.algorithm.coloring, sq.algorithm.sa_naive):
        an.set_preferences(algorithm=algo)
        assert an.get_preferences()["algorithm"] == algo
    an.set_preferences(algorithm=sq.algorithm.default)
    assert an.get_preferences()["algorithm"] == sq.algorithm.coloring
    an.set_preferences(algorithm=sq.algorithm.sa_default)
    assert an.get_preferences()["algorithm"] == sq.algorithm.sa_naive
    an.set_preferences(algorithm=sq.algorithm.naive, n_trotters=1)
    an.prepare()
    assert an.get_preferences()["algorithm"] == sq.algorithm.sa_naive
    an = new_annealer(rocm, 10, 10, dtype)
    with pytest.raises(RuntimeError):
        an.anneal_one_step(0., 0.)


def spin_energy_ref(qset, G, beta):
    """test_dense_graph_annealer.py:292-299"""
    m = len(qset)
    E = 0.
    for im in range(m):
        E += np.sum(qset[im].astype(np.float64) * qset[(im + 1) % m])
    return E * 1. / (2 * beta) * np.log(1 / np.tanh(G * beta / m))


@pytest.mark.parametrize("dtype", DTYPES)
def test_system_energy_identities(rocm, sq, dtype):
    G, beta = 0.01, 1. / 0.05
    N, m = 10, 5
    rng = np.random.default_rng(11)
    ann = rocm.dense_graph_annealer(dense_graph_random(N, SEED_BASE + 3, dtype), dtype=dtype, n_trotters=m,
                                    algorithm=sq.algorithm.sa_default)
    ann.prepare()
    ann.set_qset((2 * rng.integers(0, 2, (m, N)) - 1).astype(np.int8))
    assert np.allclose(ann.get_system_E(G, beta), np.mean(ann.get_E()), atol=epu(dtype))
    ann = rocm.dense_graph_annealer(np.ones((N, N), dtype), dtype=dtype, n_trotters=m)
    ann.prepare()
    for sign in (1, -1):
        ann.set_q(sign * np.ones(N, np.int8))
        Esys = ann.get_system_E(G, beta)
        assert np.allclose(Esys, np.mean(ann.get_E()) + spin_energy_ref(ann.get_q(), G, beta), atol=epu(dtype))
    ann = rocm.dense_graph_annealer(dense_graph_random(N, SEED_BASE + 4, dtype), dtype=dtype, n_trotters=m)
    ann.prepare()
    checker = np.fromfunction(lambda y, i: np.where((y + i) % 2 == 0, 1, -1), (m, N)).astype(np.int8)
    for qset in (checker, (2 * rng.integers(0, 2, (m, N)) - 1).astype(np.int8)):
        ann.set_qset(qset)
        Esys = ann.get_system_E(G, beta)
        assert np.allclose(Esys, np.mean(ann.get_E()) + spin_energy_ref(ann.get_q(), G, beta),
                           atol=epu(dtype) * 10)


@pytest.mark.parametrize("algo,m,opt", [("coloring", 8, 0), ("naive", 6, 0), ("coloring", 5, 1), ("sa_naive", 3, 0)])
def test_system_energy_trace_equals_oracle_per_step(oracle, rocm, sq, algo, m, opt):
    """The device trace records, after every step, exactly what get_system_E
    returns then: compared with the oracle's getSystemE
    (CPUDenseGraphAnnealer.cpp:230-247) after each step of the same
    trajectory (integer W, the MT19937 stream), and with the device's own
    get_system_E, bit for bit."""
    N = 48
    W = dense_int_W(N, SEED_BASE + 20 + m)
    ora = oracle.DenseGraphAnnealer(np.float32, 1)
    ora.seed(9)
    ora.set_qubo(W, opt)
    ora.set_m(m)
    ora.set_algorithm({"coloring": oracle.ALGO_COLORING, "naive": oracle.ALGO_NAIVE,
                       "sa_naive": oracle.ALGO_SA_NAIVE}[algo])
    ora.prepare()
    ora.randomize_spin()
    an = rocm.dense_graph_annealer(W, sq.maximize if opt else sq.minimize, dtype=np.float32, n_trotters=m,
                                   algorithm=algo, rng="mt19937")
    an.seed(9)
    an.prepare()
    an.randomize_spin()
    steps = [(G, 50.0 if algo != "sa_naive" else 0.0) for G in schedule()[:60]]
    an.trace_system_E(len(steps) - 10)
    ref = []
    for G, beta in steps:
        ora.anneal_one_step(G, beta)
        an.anneal_one_step(G, beta)
        ref.append(np.float32(ora.get_system_E(np.float32(G), np.float32(beta))))
    trace, n_steps = an.get_system_E_trace()
    assert n_steps == len(steps) and len(trace) == len(steps) - 10
    assert np.array_equal(trace.astype(np.float32), np.array(ref[:len(trace)], np.float32))
    G, beta = steps[-1]
    assert np.float32(an.get_system_E(G, beta)) == ref[-1]


class _Untraced:
    """An annealer without the trace: the adapter then calls get_system_E
    after every step or level, as the reference adapter does."""

    def __init__(self, an):
        self._an = an

    def __getattr__(self, name):
        if name in ("trace_system_E", "get_system_E_trace"):
            raise AttributeError(name)
        return getattr(self._an, name)


def test_wildqat_adapter(rocm):
    """sqaod_amd.wildqat.opt (sqaodpy/sqaod/wildqat/opt.py): sa() and sqa()
    record the energies the reference adapter records (after every SA step,
    after every SQA level) -- from the device trace, equal to get_system_E
    called per step on the same trajectory -- and return x of the right
    shape; the QUBO's optimum is found."""
    from sqaod_amd import wildqat

    class Pkg:
        traced = True

        @staticmethod
        def dense_graph_annealer(dtype):
            an = rocm.dense_graph_annealer(dtype=dtype)
            return an if Pkg.traced else _Untraced(an)

    N = 12
    Q = dense_int_W(N, SEED_BASE + 33, lo=-3, hi=3).astype(np.float64)
    results = {}
    for traced in (True, False):
        Pkg.traced = traced
        o = wildqat.opt(pkg=Pkg)
        o.qubo = Q
        o.ite = 240
        o.ann.seed(77)
        x_sa = o.sa()
        E_sa = list(o.E)
        o2 = wildqat.opt(pkg=Pkg)
        o2.qubo = Q
        o2.tro = 4
        o2.ite = 480
        o2.ann.seed(78)
        x_sqa = o2.sqa()
        E_sqa = list(o2.E)
        results[traced] = (x_sa, E_sa, x_sqa, E_sqa)
    (x_sa, E_sa, x_sqa, E_sqa), (ux_sa, uE_sa, ux_sqa, uE_sqa) = results[True], results[False]
    n_levels_sa = 0
    T = 5.0
    while T > 0.02:
        n_levels_sa += 1
        T *= 0.95
    assert len(E_sa) == n_levels_sa * ((240 + N - 1) // N)
    assert np.array_equal(np.float32(E_sa), np.float32(uE_sa)) and np.array_equal(x_sa, ux_sa)
    assert np.array_equal(np.float32(E_sqa), np.float32(uE_sqa))
    assert all(np.array_equal(a, b) for a, b in zip(x_sqa, ux_sqa)) and len(x_sqa) == 4
    assert x_sa.shape == (N,)
    # the optimum of the small QUBO (brute force)
    xs = ((np.arange(1 << N)[:, None] >> np.arange(N - 1, -1, -1)[None, :]) & 1).astype(np.float64)
    Emin = np.min(np.einsum("bi,ij,bj->b", xs, Q, xs))
    assert float(x_sa @ Q @ x_sa) == Emin
