"""Formulas objects on the rocm device (formulas.inc:586-606) against closed
forms, the oracle's Hamiltonians and the identities of the reference's
sqaodpy/tests/test_dense_graph_formulas.py and test_bipartite_graph_formulas.py.
Integer-valued inputs make every energy exact, so the checks are equalities;
for quantized random inputs (example_problems.py:16-22) the reference's own
tolerance applies."""
import numpy as np
import pytest

from problems import dense_graph_random, quantize, SEED_BASE

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64]


@pytest.fixture(scope="module")
def F():
    import sqaod_amd.rocm as rocm
    return rocm.formulas


def bits(N):
    return ((np.arange(1 << N)[:, None] >> np.arange(N - 1, -1, -1)[None, :]) & 1).astype(np.int8)


def int_sym(N, seed, lo=-4, hi=4):
    rng = np.random.default_rng(seed)
    U = rng.integers(lo, hi + 1, (N, N)).astype(np.float64)
    return np.triu(U) + np.triu(U, 1).T


@pytest.mark.parametrize("dtype", DTYPES)
def test_dense_energy_closed_forms(F, dtype):
    """test_dense_graph_formulas.py:19-51: x = 0 gives 0, x = 1 gives sum W,
    and the batch equals the single-vector function on all 2^8 x."""
    N = 8
    W = int_sym(N, SEED_BASE + 1)
    assert F.dense_graph_calculate_E(W, np.zeros(N, np.int8), dtype) == 0.0
    assert F.dense_graph_calculate_E(W, np.ones(N, np.int8), dtype) == W.sum()
    X = bits(N)
    Eb = F.dense_graph_batch_calculate_E(W, X, dtype)
    exact = np.einsum("ki,ij,kj->k", X.astype(np.float64), W, X.astype(np.float64))
    assert Eb.dtype == dtype and np.array_equal(Eb.astype(np.float64), exact)
    assert all(F.dense_graph_calculate_E(W, X[i], dtype) == exact[i] for i in range(0, 256, 17))


@pytest.mark.parametrize("dtype", DTYPES)
def test_dense_hamiltonian_and_ising_energy(oracle, F, dtype):
    """test_dense_graph_formulas.py:53-95: -sum h + sum J + c = 0, the
    Hamiltonian equals the oracle's, and E_qubo(x) = E_ising(2x - 1) on all x."""
    N = 9
    W = int_sym(N, SEED_BASE + 2)
    h, J, c = F.dense_graph_calculate_hamiltonian(W, dtype)
    rh, rJ, rc = oracle.dense_hamiltonian(W)
    assert np.array_equal(h, rh.astype(dtype)) and np.array_equal(J, rJ.astype(dtype)) and c == dtype(rc)
    assert -h.sum() + J.sum() + c == 0.0
    X = bits(N)
    Eq = F.dense_graph_batch_calculate_E(W, X, dtype)
    Ei = F.dense_graph_batch_calculate_E_from_spin(h, J, c, 2 * X - 1, dtype)
    assert np.array_equal(Eq, Ei)
    assert F.dense_graph_calculate_E_from_spin(h, J, c, 2 * X[77] - 1, dtype) == Eq[77]


@pytest.mark.parametrize("dtype", DTYPES)
def test_dense_quantized_random_within_reference_tolerance(F, dtype):
    """example_problems.py:20-22 weights: agreement with float64 numpy within
    the reference tests' tolerance (1e-6 float32 / 1e-12 float64, scaled)."""
    N = 64
    W = dense_graph_random(N, SEED_BASE + 3)
    rng = np.random.default_rng(4)
    X = rng.integers(0, 2, (50, N)).astype(np.int8)
    E = F.dense_graph_batch_calculate_E(W, X, dtype)
    ref = np.einsum("ki,ij,kj->k", X.astype(np.float64), W.astype(dtype).astype(np.float64), X.astype(np.float64))
    tol = (1e-6 if dtype == np.float32 else 1e-12) * max(1.0, np.abs(ref).max())
    assert np.max(np.abs(E - ref)) <= tol


def bg_int(N0, N1, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(-4, 5, N0).astype(np.float64), rng.integers(-4, 5, N1).astype(np.float64),
            rng.integers(-4, 5, (N1, N0)).astype(np.float64))


@pytest.mark.parametrize("dtype", DTYPES)
def test_bipartite_energy_forms(F, dtype):
    """test_bipartite_graph_formulas.py:28-80: zero / one x, the batch and the
    2-D batch equal b0.x0 + b1.x1 + x1 W x0."""
    N0, N1 = 6, 5
    b0, b1, W = bg_int(N0, N1, SEED_BASE + 5)
    z0, z1 = np.zeros(N0, np.int8), np.zeros(N1, np.int8)
    assert F.bipartite_graph_calculate_E(b0, b1, W, z0, z1, dtype) == 0.0
    o0, o1 = np.ones(N0, np.int8), np.ones(N1, np.int8)
    assert F.bipartite_graph_calculate_E(b0, b1, W, o0, o1, dtype) == b0.sum() + b1.sum() + W.sum()
    X0, X1 = bits(N0), bits(N1)
    E2 = F.bipartite_graph_batch_calculate_E_2d(b0, b1, W, X0, X1, dtype)
    exact2 = (X0 @ b0)[None, :] + (X1 @ b1)[:, None] + X1.astype(np.float64) @ W @ X0.T.astype(np.float64)
    assert E2.shape == (1 << N1, 1 << N0) and np.array_equal(E2, exact2.astype(dtype))
    idx0 = np.arange(32) % (1 << N0)
    idx1 = (np.arange(32) * 7) % (1 << N1)
    Eb = F.bipartite_graph_batch_calculate_E(b0, b1, W, X0[idx0], X1[idx1], dtype)
    assert np.array_equal(Eb, exact2[idx1, idx0].astype(dtype))


@pytest.mark.parametrize("dtype", DTYPES)
def test_bipartite_hamiltonian_and_ising_energy(oracle, F, dtype):
    """test_bipartite_graph_formulas.py:81-155: the Hamiltonian equals the
    oracle's and E_qubo(x0, x1) = E_ising(2x0 - 1, 2x1 - 1)."""
    N0, N1 = 7, 6
    b0, b1, W = bg_int(N0, N1, SEED_BASE + 6)
    h0, h1, J, c = F.bipartite_graph_calculate_hamiltonian(b0, b1, W, dtype)
    r = oracle.bipartite_hamiltonian(b0, b1, W)
    assert all(np.array_equal(a, np.asarray(b).astype(dtype)) for a, b in zip((h0, h1, J), r[:3]))
    assert c == dtype(r[3])
    rng = np.random.default_rng(7)
    X0 = rng.integers(0, 2, (40, N0)).astype(np.int8)
    X1 = rng.integers(0, 2, (40, N1)).astype(np.int8)
    Eq = F.bipartite_graph_batch_calculate_E(b0, b1, W, X0, X1, dtype)
    Ei = F.bipartite_graph_batch_calculate_E_from_spin(h0, h1, J, c, 2 * X0 - 1, 2 * X1 - 1, dtype)
    assert np.array_equal(Eq, Ei)
    assert F.bipartite_graph_calculate_E_from_spin(h0, h1, J, c, 2 * X0[3] - 1, 2 * X1[3] - 1, dtype) == Eq[3]


def test_shape_errors(F):
    # a non-square W fails in the symmetry test first, as in the reference
    # (common.py:41-62 compares W with W.T before checkers run): numpy's ValueError
    with pytest.raises(ValueError):
        F.dense_graph_calculate_E(np.ones((3, 4)), np.zeros(3, np.int8), np.float32)
    with pytest.raises(RuntimeError):
        F.dense_graph_calculate_E(np.ones((3, 3)), np.zeros(4, np.int8), np.float32)
    with pytest.raises(RuntimeError):
        F.bipartite_graph_batch_calculate_E(np.ones(3), np.ones(2), np.ones((3, 3)), np.zeros((1, 3), np.int8),
                                            np.zeros((1, 2), np.int8), np.float32)
