"""Bipartite annealer on the rocm device against the oracle
(CPUBipartiteGraphAnnealer, one worker) and the reference tests' intent
(sqaodpy/tests/test_bipartite_graph_annealer.py).

Bit-exact bar: integer-valued b0, b1, W (so J, h, dEmat are exact in float)
with rng=mt19937 -- every spin of every slice equals the oracle's after every
step (coloring with even and odd m, sa_coloring, maximize, bf16x2 and vector
GEMM forms).  Float data: the final-energy distributions over seeds match the
oracle's (two-sample KS)."""
import numpy as np
import pytest

from problems import SEED_BASE, schedule

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def int_problem(N0, N1, seed, lo=-8, hi=8, dtype=np.float32):
    rng = np.random.default_rng(seed)
    b0 = rng.integers(lo, hi + 1, N0).astype(dtype)
    b1 = rng.integers(lo, hi + 1, N1).astype(dtype)
    W = rng.integers(lo, hi + 1, (N1, N0)).astype(dtype)
    return b0, b1, W


def pair_arrays(qpairs):
    return np.stack([p[0] for p in qpairs]), np.stack([p[1] for p in qpairs])


def make_pair(rocm, oracle, b0, b1, W, m, dtype, opt=0, algo="coloring", seed=11, **prefs):
    an = rocm.bipartite_graph_annealer(b0, b1, W, optimize=opt, dtype=dtype, n_trotters=m, algorithm=algo,
                                       rng="mt19937", **prefs)
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    ref = oracle.BipartiteGraphAnnealer(dtype)
    ref.seed(seed)
    ref.set_qubo(b0, b1, W, optimize=opt)
    ref.set_m(m)
    ref.set_algorithm(7 if algo == "sa_coloring" else 3)
    ref.prepare()
    ref.randomize_spin()
    return an, ref


@pytest.mark.parametrize("N0,N1,m,algo,opt,prefs", [
    (40, 24, 8, "coloring", 0, {}),
    (33, 20, 5, "coloring", 0, {}),                # odd m: slices 0 and m-1 are both even
    (300, 100, 6, "coloring", 0, {}),              # ragged, unequal padding (N0p = 512, N1p = 256)
    (64, 48, 4, "coloring", 1, {}),                # maximize
    (50, 70, 3, "sa_coloring", 0, {}),
    (40, 24, 1, "coloring", 0, {}),                # m = 1 selects sa_coloring
    (40, 24, 7, "coloring", 0, {"field_precision": "bf16x2"}),
    (40, 24, 6, "coloring", 0, {"refresh_engine": "vector"}),
])
def test_bit_exact_against_oracle(oracle, rocm, N0, N1, m, algo, opt, prefs):
    b0, b1, W = int_problem(N0, N1, SEED_BASE + 500 + N0 + m)
    an, ref = make_pair(rocm, oracle, b0, b1, W, m, np.float32, opt=opt, algo=algo, **prefs)
    q0, q1 = pair_arrays(an.get_q())
    r0, r1 = ref.get_q()
    assert np.array_equal(q0, r0) and np.array_equal(q1, r1)
    steps = schedule()[:60] if algo == "coloring" else [5.0 * 0.9 ** k for k in range(60)]
    for k, G in enumerate(steps):
        an.anneal_one_step(G, 50.0)
        ref.anneal_one_step(G, 50.0)
        q0, q1 = pair_arrays(an.get_q())
        r0, r1 = ref.get_q()
        assert np.array_equal(q0, r0) and np.array_equal(q1, r1), f"step {k}"
    assert np.array_equal(an.get_E(), ref.calculate_E())


def test_bit_exact_float64_coloring(oracle, rocm):
    """float64 solver: exp is the device's (not glibc's) and differs from the
    host's by at most 1 ulp, so a decision can differ only when u falls within
    1 ulp of the threshold (probability ~1e-16 per spin); over this run the
    trajectories agree."""
    b0, b1, W = int_problem(30, 22, SEED_BASE + 611, dtype=np.float64)
    an, ref = make_pair(rocm, oracle, b0, b1, W, 6, np.float64)
    for G in schedule()[:30]:
        an.anneal_one_step(G, 50.0)
        ref.anneal_one_step(G, 50.0)
    q0, q1 = pair_arrays(an.get_q())
    r0, r1 = ref.get_q()
    assert np.array_equal(q0, r0) and np.array_equal(q1, r1)


def test_hamiltonian_and_energy_against_oracle(oracle, rocm):
    """test_bipartite_graph_annealer.py:64-86: get_hamiltonian equals the
    formulas, set_hamiltonian round-trips; energies equal the FP64 oracle."""
    b0, b1, W = int_problem(37, 29, SEED_BASE + 37, lo=-5, hi=5, dtype=np.float64)
    an = rocm.bipartite_graph_annealer(b0, b1, W, dtype=np.float64, n_trotters=4)
    h0, h1, J, c = an.get_hamiltonian()
    rh0, rh1, rJ, rc = oracle.bipartite_hamiltonian(b0, b1, W)
    assert np.array_equal(h0, rh0) and np.array_equal(h1, rh1) and np.array_equal(J, rJ) and c == rc
    an2 = rocm.bipartite_graph_annealer(dtype=np.float64)
    an2.set_hamiltonian(h0, h1, J, c)
    an2.prepare()
    g = an2.get_hamiltonian()
    assert all(np.array_equal(a, b) for a, b in zip(g[:3], (h0, h1, J))) and g[3] == c
    an.seed(3)
    an.prepare()
    an.randomize_spin()
    q0, q1 = pair_arrays(an.get_q())
    assert np.array_equal(an.get_E(), oracle.bipartite_energy_q(h0, h1, J, c, q0, q1))


def test_qubo_energy_all_states(rocm):
    """test_bipartite_graph_annealer.py:106-132: for every (x0, x1) of an 5 x 4
    problem, E = b0.x0 + b1.x1 + x1 W x0."""
    N0, N1 = 5, 4
    b0, b1, W = int_problem(N0, N1, SEED_BASE + 9, lo=-4, hi=4)
    an = rocm.bipartite_graph_annealer(b0, b1, W, dtype=np.float32, n_trotters=1)
    an.prepare()
    for i in range(1 << N0):
        x0 = ((i >> np.arange(N0 - 1, -1, -1)) & 1).astype(np.float32)
        for j in range(1 << N1):
            x1 = ((j >> np.arange(N1 - 1, -1, -1)) & 1).astype(np.float32)
            an.set_q(((2 * x0 - 1).astype(np.int8), (2 * x1 - 1).astype(np.int8)))
            assert an.get_E()[0] == b0 @ x0 + b1 @ x1 + x1 @ W @ x0


def test_known_answers_and_state_api(rocm):
    """test_bipartite_graph_annealer.py:88-104, 134-194: E(q=-1) = 0, E(ones)
    = N0 N1 + N0 + N1, set_q / set_qset / get_x / problem size / n_trotters."""
    N0, N1 = 4, 4
    an = rocm.bipartite_graph_annealer(np.ones(N0), np.ones(N1), np.ones((N1, N0)), dtype=np.float32, n_trotters=1)
    an.prepare()
    an.set_q((np.ones(N0, np.int8), np.ones(N1, np.int8)))
    assert an.get_E().tolist() == [N0 * N1 + N0 + N1]
    x = an.get_x()
    assert len(x) == 1 and np.all(x[0][0] == 1) and np.all(x[0][1] == 1)
    an.set_q((-np.ones(N0, np.int8), -np.ones(N1, np.int8)))
    assert an.get_E().tolist() == [0.0]
    b0, b1, W = int_problem(100, 50, 5)
    an = rocm.bipartite_graph_annealer(b0, b1, W, dtype=np.float32, n_trotters=150)
    assert an.get_problem_size() == (100, 50)
    an.prepare()
    rng = np.random.default_rng(1)
    q0in = (2 * rng.integers(0, 2, 100) - 1).astype(np.int8)
    q1in = (2 * rng.integers(0, 2, 50) - 1).astype(np.int8)
    an.set_q((q0in, q1in))
    assert all(np.array_equal(p[0], q0in) and np.array_equal(p[1], q1in) for p in an.get_q())
    qset = [((2 * rng.integers(0, 2, 100) - 1).astype(np.int8), (2 * rng.integers(0, 2, 50) - 1).astype(np.int8))
            for _ in range(37)]
    an.set_qset(qset)
    out = an.get_q()
    assert len(out) == 37 and an.get_preferences()['n_trotters'] == 37
    assert all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) for a, b in zip(qset, out))
    an.set_preferences(n_trotters=3)
    assert an.get_preferences()['n_trotters'] == 3


@pytest.mark.parametrize("algo,m", [("coloring", 6), ("default", 6), ("naive", 6), ("sa_coloring", 1), ("sa_naive", 1)])
def test_trivial_problems_reach_zero(oracle, rocm, algo, m):
    """test_bipartite_graph_annealer.py:196-266 (minimize, maximize,
    set_hamiltonian), with the CUDA device's algorithm fallbacks."""
    import sqaod_amd as sq
    N0, N1 = 10, 8
    sa = algo.startswith("sa")
    Gi, Gf, beta = (10.0, 0.02, 1.0) if sa else (5.0, 0.02, 1 / 0.03)
    for case in ("min", "max", "ham"):
        sgn = -1.0 if case == "max" else 1.0
        b0, b1, W = sgn * np.ones(N0), sgn * np.ones(N1), sgn * np.ones((N1, N0))
        an = rocm.bipartite_graph_annealer(dtype=np.float32)
        if case == "ham":
            an.set_hamiltonian(*oracle.bipartite_hamiltonian(b0, b1, W))
        else:
            an.set_qubo(b0, b1, W, sq.maximize if case == "max" else sq.minimize)
        an.set_preferences(n_trotters=m, algorithm=algo)
        assert an.get_preferences()['algorithm'] == ("sa_coloring" if sa else "coloring")
        an.seed(21)
        an.prepare()
        an.randomize_spin()
        G, tau = Gi, np.exp(np.log(Gf / Gi) / 100)
        for _ in range(100):
            an.anneal_one_step(G, beta)
            G *= tau
        E = an.get_E()
        assert (E.max() if case == "max" else E.min()) == 0.0


def test_system_energy_matches_oracle(oracle, rocm):
    """getSystemE (:268-286): mean E - spin overlap * (0.5/beta) ln tanh(G beta/m)."""
    b0, b1, W = int_problem(20, 16, SEED_BASE + 20, dtype=np.float64)
    an, ref = make_pair(rocm, oracle, b0, b1, W, 5, np.float64)
    for G in (3.0, 1.0, 0.3):
        an.anneal_one_step(G, 10.0)
        ref.anneal_one_step(G, 10.0)
    assert an.get_system_E(0.5, 10.0) == pytest.approx(ref.get_system_E(0.5, 10.0), rel=1e-12, abs=1e-9)


def test_float_problem_energy_distribution_matches_oracle(oracle, rocm):
    """b0, b1, W ~ N(0,1), N0 = N1 = 96, m = 16, 60 steps: the final mean
    energies over 24 seeds (Philox streams on the device, MT19937 on the CPU)
    are indistinguishable by a two-sample KS test (p > 0.001)."""
    from scipy.stats import ks_2samp
    N0 = N1 = 96
    m = 16
    rng = np.random.default_rng(SEED_BASE + 2)
    b0 = rng.standard_normal(N0).astype(np.float32)
    b1 = rng.standard_normal(N1).astype(np.float32)
    W = rng.standard_normal((N1, N0)).astype(np.float32)
    steps = schedule(G0=5.0, Gmin=0.05, tau=0.92)
    gpu, cpu = [], []
    for s in range(24):
        an = rocm.bipartite_graph_annealer(b0, b1, W, dtype=np.float32, n_trotters=m)
        an.seed(1000 + s)
        an.prepare()
        an.randomize_spin()
        ref = oracle.BipartiteGraphAnnealer(np.float32)
        ref.seed(1000 + s)
        ref.set_qubo(b0, b1, W)
        ref.set_m(m)
        ref.prepare()
        ref.randomize_spin()
        for G in steps:
            an.anneal_one_step(G, 20.0)
            ref.anneal_one_step(G, 20.0)
        gpu.append(float(np.mean(an.get_E())))
        cpu.append(float(np.mean(ref.calculate_E())))
    assert ks_2samp(gpu, cpu).pvalue > 1e-3, (np.mean(gpu), np.mean(cpu))
