"""Bipartite brute-force searcher on the rocm device against the oracle's
restatement of the one-worker CPU search (CPUBipartiteGraphBFSearcher,
CPUBipartiteGraphBatchSearch) and the reference tests' known answers
(sqaodpy/tests/test_bipartite_graph_bf_searcher.py).  Integer data make every
energy exact, so minima and pair lists must be equal."""
import numpy as np
import pytest

from problems import SEED_BASE

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def packed(x):
    x = np.asarray(x, np.uint64)
    N = x.shape[-1]
    return (x << np.arange(N - 1, -1, -1, dtype=np.uint64)).sum(axis=-1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_known_answers(rocm, dtype):
    """test_bipartite_graph_bf_searcher.py:41-110."""
    s = rocm.bipartite_graph_bf_searcher(np.ones(4), np.ones(4), np.ones((4, 4)), optimize=1, dtype=dtype)
    s.search()
    E = s.get_E()
    assert len(E) == 1 and E[0] == 4 * 4 + 4 + 4
    x = s.get_x()
    assert len(x) == 1 and np.all(x[0][0] == 1) and np.all(x[0][1] == 1)
    assert s.get_problem_size() == (4, 4)
    for opt, sign, Eexp, bit in [(0, 1, 0, 0), (0, -1, -80, 1), (1, 1, 80, 1), (1, -1, 0, 0)]:
        s.set_qubo(sign * np.ones(8), sign * np.ones(8), sign * np.ones((8, 8)), opt)
        s.search()
        E, x = s.get_E(), s.get_x()
        assert E[0] == Eexp and len(x) == 1
        assert np.all(x[0][0] == bit) and np.all(x[0][1] == bit)


@pytest.mark.parametrize("N0,N1,T0,opt,seed", [(5, 4, 1024, 0, 1), (3, 7, 256, 1, 2), (10, 6, 256, 0, 3),
                                               (12, 8, 512, 1, 4), (14, 9, 1024, 0, 5), (10, 13, 256, 0, 6)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_matches_oracle(oracle, rocm, N0, N1, T0, opt, seed, dtype):
    """b0, b1, W integers in [-2, 2] (ties are common): Emin and the pair list
    equal the oracle's -- the first N0 + N1 minimisers in the order (x0 tile,
    x1, x0 in tile) -- including x0 tiles smaller than 2^N0 and searches
    that take several scan rectangles (N1 = 13: two x1 chunks)."""
    rng = np.random.default_rng(SEED_BASE + seed)
    b0 = rng.integers(-2, 3, N0).astype(np.float64)
    b1 = rng.integers(-2, 3, N1).astype(np.float64)
    W = rng.integers(-2, 3, (N1, N0)).astype(np.float64)
    Eo, xs0, xs1 = oracle.bipartite_bf_search(b0, b1, W, dtype=dtype, optimize=opt, tile_size0=T0)
    s = rocm.bipartite_graph_bf_searcher(b0, b1, W, optimize=opt, dtype=dtype, tile_size_0=T0)
    s.prepare()
    calls = 0
    while True:
        done, cx0, cx1 = s.search_range()
        calls += 1
        if done:
            break
    assert cx0 == 1 << N0 and cx1 == 0
    if N1 == 13:
        assert calls >= 2
    s.make_solution()
    E = s.get_E()
    x = s.get_x()
    assert E.dtype == dtype and np.all(E == Eo) and len(x) == len(xs0)
    assert np.array_equal(packed([p[0] for p in x]), xs0)
    assert np.array_equal(packed([p[1] for p in x]), xs1)


def test_all_ties_keep_first_pairs_in_visiting_order(oracle, rocm):
    N0, N1 = 9, 5
    z0, z1, zW = np.zeros(N0), np.zeros(N1), np.zeros((N1, N0))
    for T0 in (256, 512):
        Eo, xs0, xs1 = oracle.bipartite_bf_search(z0, z1, zW, tile_size0=T0)
        s = rocm.bipartite_graph_bf_searcher(z0, z1, zW, dtype=np.float32, tile_size_0=T0)
        s.search()
        x = s.get_x()
        assert len(x) == N0 + N1
        assert np.array_equal(packed([p[0] for p in x]), xs0)
        assert np.array_equal(packed([p[1] for p in x]), xs1)


def test_errors(rocm):
    s = rocm.bipartite_graph_bf_searcher(dtype=np.float32)
    with pytest.raises(RuntimeError):
        s.set_qubo(np.ones(64), np.ones(2), np.ones((2, 64)))
    with pytest.raises(RuntimeError):
        s.set_qubo(np.ones(3), np.ones(2), np.ones((3, 2)))
    with pytest.raises(RuntimeError):
        s.prepare()
