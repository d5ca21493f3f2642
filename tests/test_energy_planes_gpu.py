"""Annealer energies on the integer MFMA units (calculate_E / make_solution,
DGFuncs::calculate_E at sqaodc/cpu/SharedFormulas.cpp:69-83 and
BGFuncs::calculate_E at :184-202), read back in float64 with get_E_fp64.

The device writes J as int8 digit planes scaled by a power of two and sums
the plane products exactly in int32, so:
* on integer problems (J multiples of 1/4, several planes when the weights
  are large) every energy equals the exact value, computed here in int64
  from the annealer's own h, J, c;
* on N(0, 1) weights the energy is within 1e-12 relative of a host FP64
  evaluation of the same h, J, c (float32 J is held exactly; float64 J is
  rounded at 2^-54 of its largest element);
* two evaluations of the same spins give the same bits.
"""
import numpy as np
import pytest

from problems import SEED_BASE, dense_int_W, dense_normal_W, schedule

pytestmark = pytest.mark.gpu

BETA = 50.0


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def dense_exact_E(h, J, c, Q):
    """-c - q.h - q^T J q exactly: 4 h, 4 J, 4 c are integers (J = -W/4 of
    integer W; h and c are roundings of multiples of 1/4)."""
    h4 = np.asarray(h, np.float64) * 4
    J4 = np.asarray(J, np.float64) * 4
    c4 = scalar(c) * 4
    assert np.all(h4 == np.rint(h4)) and np.all(J4 == np.rint(J4)) and c4 == np.rint(c4)
    Qi = np.asarray(Q, np.int64)
    F = Qi @ J4.astype(np.int64)  # exact: |4 J| <= 2^20, N <= 2^11
    E4 = -np.int64(c4) - Qi @ h4.astype(np.int64) - np.einsum("yi,yi->y", F, Qi)
    return E4.astype(np.float64) / 4.0


def scalar(c):
    return float(np.asarray(c, np.float64).reshape(-1)[0])


def dense_fp64_E(h, J, c, Q):
    Q = np.asarray(Q, np.float64)
    return -scalar(c) - Q @ np.asarray(h, np.float64) - np.einsum("yi,yi->y", Q @ np.asarray(J, np.float64), Q)


def dense_solution(rocm, W, m, dtype, steps, seed=1000, **prefs):
    an = rocm.dense_graph_annealer(W.astype(dtype), dtype=dtype, n_trotters=m, **prefs)
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    for G in schedule()[:steps]:
        an.anneal_one_step(G, BETA)
    an.make_solution()
    return an


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("N,m,wmax", [(1000, 12, 1 << 20), (128, 8, 8), (17, 3, 1000), (1, 2, 5), (65, 1, 1 << 12)])
def test_dense_integer_energies_exact(rocm, dtype, N, m, wmax):
    W = dense_int_W(N, SEED_BASE + N, lo=-wmax, hi=wmax, dtype=np.float64)
    an = dense_solution(rocm, W, m, dtype, steps=5)
    h, J, c = an.get_hamiltonian()
    Q = np.stack(an.get_q())
    E = an.get_E_fp64()
    assert np.array_equal(E, dense_exact_E(h, J, c, Q))
    assert np.array_equal(an.get_E(), E.astype(dtype))


def test_dense_zero_problem(rocm):
    an = dense_solution(rocm, np.zeros((40, 40)), 4, np.float32, steps=2)
    assert np.array_equal(an.get_E_fp64(), np.zeros(4))


@pytest.mark.parametrize("dtype,N,m", [(np.float32, 4096, 256), (np.float64, 2048, 64)])
def test_dense_normal_energies_fp64(rocm, dtype, N, m):
    W = dense_normal_W(N, SEED_BASE, dtype=np.float64)
    an = dense_solution(rocm, W, m, dtype, steps=3)
    h, J, c = an.get_hamiltonian()
    Q = np.stack(an.get_q())
    E = an.get_E_fp64()
    ref = dense_fp64_E(h, J, c, Q)
    rel = np.max(np.abs(E - ref) / np.abs(ref))
    assert rel <= 1e-12, rel
    # deterministic: the same spins give the same bits
    an.calculate_E()
    assert np.array_equal(an.get_E_fp64(), E)


def bipartite_exact_E(h0, h1, J, c, Q0, Q1):
    h04, h14, J4 = (np.asarray(a, np.float64) * 4 for a in (h0, h1, J))
    c4 = scalar(c) * 4
    for a in (h04, h14, J4, c4):
        assert np.all(a == np.rint(a))
    Q0i, Q1i = np.asarray(Q0, np.int64), np.asarray(Q1, np.int64)
    E4 = (-np.int64(c4) - Q0i @ h04.astype(np.int64) - Q1i @ h14.astype(np.int64)
          - np.einsum("yi,yi->y", Q1i, Q0i @ J4.astype(np.int64).T))
    return E4.astype(np.float64) / 4.0


def bipartite_solution(rocm, b0, b1, W, m, dtype, steps, seed=1000):
    an = rocm.bipartite_graph_annealer(b0.astype(dtype), b1.astype(dtype), W.astype(dtype), dtype=dtype,
                                       n_trotters=m)
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    for G in schedule()[:steps]:
        an.anneal_one_step(G, BETA)
    an.make_solution()
    return an


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("N0,N1,m,wmax", [(300, 100, 6, 1 << 18), (64, 64, 4, 8), (5, 70, 3, 100)])
def test_bipartite_integer_energies_exact(rocm, dtype, N0, N1, m, wmax):
    rng = np.random.default_rng(SEED_BASE + N0 + N1)
    b0 = rng.integers(-wmax, wmax + 1, N0).astype(np.float64)
    b1 = rng.integers(-wmax, wmax + 1, N1).astype(np.float64)
    W = rng.integers(-wmax, wmax + 1, (N1, N0)).astype(np.float64)
    an = bipartite_solution(rocm, b0, b1, W, m, dtype, steps=5)
    h0, h1, J, c = an.get_hamiltonian()
    q = an.get_q()
    Q0, Q1 = np.stack([p[0] for p in q]), np.stack([p[1] for p in q])
    E = an.get_E_fp64()
    assert np.array_equal(E, bipartite_exact_E(h0, h1, J, c, Q0, Q1))


def test_bipartite_normal_energies_fp64(rocm):
    N0 = N1 = 1024
    m = 64
    rng = np.random.default_rng(SEED_BASE)
    b0, b1 = rng.standard_normal(N0), rng.standard_normal(N1)
    W = rng.standard_normal((N1, N0))
    an = bipartite_solution(rocm, b0, b1, W, m, np.float32, steps=3)
    h0, h1, J, c = an.get_hamiltonian()
    q = an.get_q()
    Q0 = np.stack([p[0] for p in q]).astype(np.float64)
    Q1 = np.stack([p[1] for p in q]).astype(np.float64)
    ref = (-scalar(c) - Q0 @ np.asarray(h0, np.float64) - Q1 @ np.asarray(h1, np.float64)
           - np.einsum("yi,yi->y", Q1, Q0 @ np.asarray(J, np.float64).T))
    E = an.get_E_fp64()
    assert np.max(np.abs(E - ref) / np.abs(ref)) <= 1e-12
    an.calculate_E()
    assert np.array_equal(an.get_E_fp64(), E)
