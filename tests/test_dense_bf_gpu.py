"""Dense brute-force searcher on the rocm device against the oracle
(CPUDenseGraphBFSearcher one-worker semantics), the reference tests'
known answers (sqaodpy/tests/test_dense_graph_bf_searcher.py:36-183) and
the CPU-vs-CUDA x-list identity checks (sqaodc/tests/
CUDADenseGraphBFSolverTest.cpp:187-247, testMatSymmetric at N=16 and 20)."""
import numpy as np
import pytest

from problems import dense_int_W, dense_normal_W, dense_graph_8x8, mat_symmetric_cycle, SEED_BASE

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def search(rocm, W, dtype=np.float32, optimize=0, **prefs):
    s = rocm.dense_graph_bf_searcher(W, optimize=optimize, dtype=dtype, **prefs)
    s.search()
    return s, s.get_E(), np.array(s.get_x(), np.int8).reshape(-1, W.shape[0])


def packed(x):
    """int8 bit rows -> packed x (bit N-1-pos = x[pos])."""
    N = x.shape[1]
    w = (1 << np.arange(N - 1, -1, -1, dtype=np.uint64))
    return (x.astype(np.uint64) * w).sum(axis=1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("N", [1, 5, 8, 12, 13, 16, 20, 24])
def test_integer_problems_match_oracle(oracle, rocm, dtype, N):
    """Integer W in [-8, 8]: Emin and the whole x list equal the oracle's."""
    W = dense_int_W(N, SEED_BASE + 400 + N).astype(dtype)
    _, E, x = search(rocm, W, dtype)
    Eo, xo = oracle.dense_bf_search(W.astype(np.float64), dtype=dtype)
    assert E.dtype == dtype
    assert float(E[0]) == Eo and np.all(E == E[0])
    assert packed(x).tolist() == xo.tolist()


@pytest.mark.parametrize("N", [16, 20])
def test_mat_symmetric_x_list(oracle, rocm, N):
    """testMatSymmetric (-2..2 cycle): equal E and identical x lists, as
    CUDADenseGraphBFSolverTest.cpp:218-247 requires of CPU and CUDA."""
    W = mat_symmetric_cycle(N, np.float32)
    for opt in (0, 1):
        _, E, x = search(rocm, W, np.float32, optimize=opt)
        Eo, xo = oracle.dense_bf_search(W.astype(np.float64), dtype=np.float32, optimize=opt)
        assert float(E[0]) == Eo
        assert packed(x).tolist() == xo.tolist()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("opt,sign,Eexp,xbit", [(0, 1, 0.0, 0), (0, -1, -64.0, 1), (1, 1, 64.0, 1), (1, -1, 0.0, 0)])
def test_ones_known_answers(rocm, dtype, opt, sign, Eexp, xbit):
    """test_dense_graph_bf_searcher.py:82-97."""
    _, E, x = search(rocm, sign * np.ones((8, 8), dtype), dtype, optimize=opt)
    assert len(E) == 1 and E[0] == Eexp
    assert x.shape == (1, 8) and np.all(x == xbit)


def test_get_E_and_x_of_all_ones_maximize(rocm):
    """test_dense_graph_bf_searcher.py:40-63: N=4, W=ones, maximize."""
    _, E, x = search(rocm, np.ones((4, 4), np.float64), np.float64, optimize=1)
    assert E.tolist() == [16.0] and x.tolist() == [[1, 1, 1, 1]]


def test_8x8_example(rocm):
    """example_problems.py:4-14: the minimum -80 is reached by every x with 4
    or 5 bits set (70 + 56 of them), listed in ascending order."""
    _, E, x = search(rocm, dense_graph_8x8(np.float32), np.float32)
    bits = x.sum(axis=1)
    assert np.all(E == -80.0) and len(E) == 126
    assert set(bits.tolist()) == {4, 5}
    p = packed(x)
    assert np.all(np.diff(p.astype(np.int64)) > 0)


@pytest.mark.parametrize("N,tile", [(12, 256), (14, 1024), (18, 512)])
def test_ties_truncated_to_tile_size(rocm, N, tile):
    """W = 0: every x ties; the list is the first tile_size x in order
    (CPUDenseGraphBatchSearch.cpp:39-41, CPUDenseGraphBFSearcher.cpp:127-133)."""
    _, E, x = search(rocm, np.zeros((N, N), np.float32), np.float32, tile_size=tile)
    assert np.all(E == 0.0) and len(E) == tile
    assert packed(x).tolist() == list(range(tile))


def test_ties_spanning_blocks_and_ranges(oracle, rocm):
    """A problem whose minimisers lie in many blocks and in every search span:
    only the last bit matters (W[N-1][N-1] = -1), so every odd x is optimal;
    the list is the first tile_size odd x, across block boundaries."""
    N = 27
    W = np.zeros((N, N), np.float32)
    W[N - 1, N - 1] = -1.0
    s, E, x = search(rocm, W, np.float32, tile_size=70000)
    p = packed(x)
    assert np.all(E == -1.0) and len(p) == 70144  # tile size rounded up to 256
    assert p.tolist() == list(range(1, 2 * 70144, 2))


def test_float_problem_close_to_oracle(oracle, rocm):
    """W ~ N(0,1), N=22: the minimum agrees within fp32 rounding and the
    minimiser is the oracle's."""
    W = dense_normal_W(22, SEED_BASE + 77)
    _, E, x = search(rocm, W, np.float32)
    Eo, xo = oracle.dense_bf_search(W.astype(np.float64), dtype=np.float32)
    assert abs(float(E[0]) - Eo) <= 1e-5 * abs(Eo)
    assert packed(x)[0] == xo[0]


def test_search_range_steps(oracle, rocm):
    """search_range advances by spans of 2^26 and reports completion; the
    result equals search()'s."""
    N = 28
    W = dense_int_W(N, SEED_BASE + 28, lo=-2, hi=2)
    s = rocm.dense_graph_bf_searcher(W, dtype=np.float32)
    s.prepare()
    ends = []
    while True:
        done, cur = s.search_range()
        ends.append(cur)
        if done:
            break
    assert ends == [1 << 26, 2 << 26, 3 << 26, 4 << 26]
    s.make_solution()
    Eo, xo = oracle.dense_bf_search(W.astype(np.float64), dtype=np.float32)
    assert float(s.get_E()[0]) == Eo
    assert packed(np.array(s.get_x(), np.int8)).tolist() == xo.tolist()


def test_calling_sequence_and_prefs(rocm):
    """test_dense_graph_bf_searcher.py:21-35, 124-153."""
    W = dense_int_W(10, SEED_BASE + 10)
    s = rocm.dense_graph_bf_searcher(W, dtype=np.float32)
    s.prepare()
    s.search_range()
    s.make_solution()
    s.search()
    s.calculate_E()
    s.get_E()
    assert s.get_problem_size() == 10
    x = s.get_x()
    assert len(x) >= 1
    prefs = s.get_preferences()
    assert prefs['algorithm'] == 'brute_force_search'
    assert prefs['device'] == 'rocm' and prefs['precision'] == 'float'
    s.set_preferences(algorithm='coloring')
    assert s.get_preferences()['algorithm'] == 'brute_force_search'


def test_problem_size_limits(rocm):
    """test_dense_graph_bf_searcher.py:64-68, 118-122: N=63 is accepted, N=100
    raises RuntimeError."""
    W = dense_int_W(63, SEED_BASE + 63)
    s = rocm.dense_graph_bf_searcher(W, dtype=np.float32)
    assert s.get_problem_size() == 63
    with pytest.raises(RuntimeError):
        rocm.dense_graph_bf_searcher(dense_int_W(100, SEED_BASE + 100), dtype=np.float32)
