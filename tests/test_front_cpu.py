"""The Python solver front ends (sqaod_amd/common/front.py) against a
recording stand-in for the device extension: the arguments each method
passes on (dtype casts of G and beta as in
sqaodpy/sqaod/common/dense_graph_annealer_base.py:95-101, symmetrized W,
int8 spins, merged preferences), the results it hands back, and the
brute-force search loop (prepare, search_range until done, make_solution;
sqaodc/common/Solver.cpp:214-218)."""
import numpy as np
import pytest

from sqaod_amd.common import front, common
from sqaod_amd.common.preference import minimize, maximize


class Recorder:
    """Extension stand-in: records (name, args) and answers a few queries."""

    def __init__(self, problem_size=3, ranges=3):
        self.calls = []
        self.problem_size = problem_size
        self.ranges_left = ranges

    def __getattr__(self, name):
        def f(*args):
            self.calls.append((name, args))
            if name == "get_problem_size":
                return self.problem_size
            if name == "search_range":
                self.ranges_left -= 1
                return (self.ranges_left <= 0, 0)
            if name == "get_hamiltonian":
                h, J, c = args[1:4]
                h[:] = 1
                J[:] = 2
                c[0] = 3
            return name
        return f


class Dense(front.DenseGraphAnnealerBase):
    def __init__(self, cext, W, optimize, dtype, prefs):
        self._cobj = "handle"
        front.DenseGraphAnnealerBase.__init__(self, cext, dtype, (W,), optimize, prefs)


def test_dense_annealer_front_passes_converted_arguments():
    rec = Recorder()
    W = np.triu(np.arange(9.0).reshape(3, 3))  # triangular: folded into (W + W^T) / 2
    a = Dense(rec, W, maximize, np.float32, {"n_trotters": 4})
    name, args = rec.calls[0]
    assert name == "set_qubo" and args[0] == "handle" and args[-1] == np.float32
    Wsent = args[1]
    assert Wsent.dtype == np.float32 and np.array_equal(Wsent, Wsent.T)
    assert np.array_equal(Wsent, ((W + W.T) * 0.5).astype(np.float32))
    assert args[2] is maximize and a.get_optimize_dir() is maximize
    assert rec.calls[1] == ("set_preferences", ("handle", {"n_trotters": 4}, np.float32))
    a.set_preferences({"rng": "mt19937"}, sweep="plain")
    assert rec.calls[-1][1][1] == {"rng": "mt19937", "sweep": "plain"}
    a.anneal_one_step(0.1, 50)
    n, args = rec.calls[-1]
    assert n == "anneal_one_step" and type(args[1]) is np.float32 and type(args[2]) is np.float32
    assert args[1] == np.float32(0.1)
    assert a.get_system_E(1.5, 2) == "get_system_E"
    a.set_q([1, -1, 1])
    assert rec.calls[-1][1][1].dtype == np.int8
    a.set_qset([[1, -1, 1], [-1, -1, 1]])
    assert all(q.dtype == np.int8 for q in rec.calls[-1][1][1])
    a.set_hamiltonian([1, 2, 3], np.eye(3), 0.5)
    n, args = rec.calls[-1]
    assert n == "set_hamiltonian" and args[3] == np.float32(0.5) and a.get_optimize_dir() is minimize
    h, J, c = a.get_hamiltonian()
    assert h.shape == (3,) and J.shape == (3, 3) and c == 3 and np.all(J == 2)
    for m in ("seed", "randomize_spin", "prepare", "calculate_E", "make_solution", "get_E", "get_x", "get_q"):
        args = (7,) if m == "seed" else ()
        assert getattr(a, m)(*args) == m
        assert rec.calls[-1] == (m, ("handle", *args, np.float32))
    with pytest.raises(RuntimeError, match="square"):
        a.set_qubo(np.zeros((2, 3)))
    with pytest.raises(RuntimeError, match="triangular nor symmetric"):
        a.set_qubo(np.array([[0.0, 1.0], [2.0, 0.0]]))
    del a
    assert rec.calls[-1] == ("delete", ("handle", np.float32))


def test_bipartite_front_shapes_and_pairs():
    rec = Recorder(problem_size=(2, 3))

    class Bip(front.BipartiteGraphAnnealerBase):
        def __init__(self):
            self._cobj = "h"
            front.BipartiteGraphAnnealerBase.__init__(self, rec, np.float64, (None, None, None), minimize, None)

    b = Bip()
    assert rec.calls == [("set_preferences", ("h", {}, np.float64))]
    with pytest.raises(RuntimeError, match="N1 x N0"):
        b.set_qubo(np.zeros(2), np.zeros(3), np.zeros((2, 3)))
    b.set_qubo([1, 2], [3, 4, 5], np.ones((3, 2)))
    args = rec.calls[-1][1]
    assert [a.shape for a in args[1:4]] == [(2,), (3,), (3, 2)] and all(a.dtype == np.float64 for a in args[1:4])
    b.set_q(([1, -1], [1, 1, -1]))
    pair = rec.calls[-1][1][1]
    assert isinstance(pair, tuple) and pair[0].dtype == np.int8 and pair[1].shape == (3,)
    h0, h1, J, c = b.get_hamiltonian()
    assert (h0.shape, h1.shape, J.shape) == ((2,), (3,), (3, 2))


def test_bf_search_loop():
    rec = Recorder(ranges=4)

    class BF(front.DenseGraphBFSearcherBase):
        def __init__(self):
            self._cobj = "s"
            front.DenseGraphBFSearcherBase.__init__(self, rec, np.float32, (np.eye(2),), minimize, {})

    s = BF()
    rec.calls.clear()
    s.search()
    assert [c[0] for c in rec.calls] == ["prepare"] + ["search_range"] * 4 + ["make_solution"]
    s._search()
    assert rec.calls[-1] == ("search", ("s", np.float32))


def test_common_helpers():
    W = np.array([[1.0, 2.0], [2.0 + 1e-9, 3.0]])
    S = common.symmetrize(W)
    assert np.array_equal(S, S.T) and S[1, 0] == 2.0
    R = common.generate_random_symmetric_W(5, -1.0, 1.0, np.float32)
    assert R.dtype == np.float32 and np.array_equal(R, R.T) and R.min() >= -1 and R.max() < 1
    bits = common.create_bitset_sequence([5, 0, 7], 3)
    assert bits.tolist() == [[1, 0, 1], [0, 0, 0], [1, 1, 1]]
    assert common.bit_to_spin([0, 1]).tolist() == [-1, 1]
    assert common.spin_to_bit([-1, 1]).tolist() == [0, 1]
    assert common.fix_type(np.arange(3), np.float32).dtype == np.float32
    pair = common.fix_type([np.arange(2), [1.5]], np.float64)  # a list: each element fixed
    assert [p.dtype for p in pair] == [np.float64, np.float64] and pair[1].tolist() == [1.5]
