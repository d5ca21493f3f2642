"""Device-independent front end of the formulas modules
(sqaodpy/sqaod/common/formulas_base.py): type fixing, symmetrization and
shape checks, then the device module's function."""
import numpy as np

from . import common


def _as_i8(x):
    return np.ascontiguousarray(np.asarray(x), np.int8)


def _batch(x):
    x = _as_i8(x)
    return x.reshape(1, -1) if x.ndim == 1 else x


def _check_qubo(W):
    if W.ndim != 2 or W.shape[0] != W.shape[1]:
        raise RuntimeError("W must be a square matrix.")


def _check_bipartite(b0, b1, W):
    if W.ndim != 2 or b0.ndim != 1 or b1.ndim != 1 or W.shape != (b1.shape[0], b0.shape[0]):
        raise RuntimeError("Shape mismatch: W must be N1 x N0 for b0[N0], b1[N1].")


# dense graph

def dense_graph_calculate_E(cext, cobj, W, x, dtype):
    W = common.symmetrize(common.fix_type(W, dtype))
    _check_qubo(W)
    x = _as_i8(x)
    if x.shape != (W.shape[0],):
        raise RuntimeError("x must be a vector of length N.")
    E = np.ndarray((1), dtype)
    cext.dense_graph_calculate_E(cobj, E, W, x, dtype)
    return E[0]


def dense_graph_batch_calculate_E(cext, cobj, W, x, dtype):
    W = common.symmetrize(common.fix_type(W, dtype))
    _check_qubo(W)
    x = _batch(x)
    if x.shape[1] != W.shape[0]:
        raise RuntimeError("x rows must have length N.")
    E = np.empty((x.shape[0]), dtype)
    cext.dense_graph_batch_calculate_E(cobj, E, W, x, dtype)
    return E


def dense_graph_calculate_hamiltonian(cext, cobj, W, dtype):
    W = common.symmetrize(common.fix_type(W, dtype))
    _check_qubo(W)
    N = W.shape[0]
    h = np.empty((N), dtype)
    J = np.empty((N, N), dtype)
    c = np.empty((1), dtype)
    cext.dense_graph_calculate_hamiltonian(cobj, h, J, c, W, dtype)
    return h, J, c[0]


def dense_graph_calculate_E_from_spin(cext, cobj, h, J, c, q, dtype):
    h, J = common.fix_type([h, J], dtype)
    J = common.symmetrize(J)
    q = _as_i8(q)
    if h.shape != (J.shape[0],) or q.shape != h.shape:
        raise RuntimeError("Shape mismatch between h, J and q.")
    E = np.ndarray((1), dtype)
    cext.dense_graph_calculate_E_from_spin(cobj, E, h, J, dtype(c), q, dtype)
    return E[0]


def dense_graph_batch_calculate_E_from_spin(cext, cobj, h, J, c, q, dtype):
    h, J = common.fix_type([h, J], dtype)
    J = common.symmetrize(J)
    q = _batch(q)
    if h.shape != (J.shape[0],) or q.shape[1] != h.shape[0]:
        raise RuntimeError("Shape mismatch between h, J and q.")
    E = np.empty((q.shape[0]), dtype)
    cext.dense_graph_batch_calculate_E_from_spin(cobj, E, h, J, dtype(c), q, dtype)
    return E


# bipartite graph

def bipartite_graph_calculate_E(cext, cobj, b0, b1, W, x0, x1, dtype):
    b0, b1, W = common.fix_type([b0, b1, W], dtype)
    _check_bipartite(b0, b1, W)
    x0, x1 = _as_i8(x0), _as_i8(x1)
    E = np.ndarray((1), dtype)
    cext.bipartite_graph_calculate_E(cobj, E, b0, b1, W, x0, x1, dtype)
    return E[0]


def bipartite_graph_batch_calculate_E(cext, cobj, b0, b1, W, x0, x1, dtype):
    b0, b1, W = common.fix_type([b0, b1, W], dtype)
    _check_bipartite(b0, b1, W)
    x0, x1 = _batch(x0), _batch(x1)
    if x0.shape[0] != x1.shape[0]:
        raise RuntimeError("Different batch dims between x0 and x1.")
    E = np.empty((x0.shape[0]), dtype)
    cext.bipartite_graph_batch_calculate_E(cobj, E, b0, b1, W, x0, x1, dtype)
    return E


def bipartite_graph_batch_calculate_E_2d(cext, cobj, b0, b1, W, x0, x1, dtype):
    b0, b1, W = common.fix_type([b0, b1, W], dtype)
    _check_bipartite(b0, b1, W)
    x0, x1 = _batch(x0), _batch(x1)
    E = np.empty((x1.shape[0], x0.shape[0]), dtype)
    cext.bipartite_graph_batch_calculate_E_2d(cobj, E, b0, b1, W, x0, x1, dtype)
    return E


def bipartite_graph_calculate_hamiltonian(cext, cobj, b0, b1, W, dtype):
    b0, b1, W = common.fix_type([b0, b1, W], dtype)
    _check_bipartite(b0, b1, W)
    N0, N1 = W.shape[1], W.shape[0]
    h0 = np.empty((N0), dtype)
    h1 = np.empty((N1), dtype)
    J = np.empty((N1, N0), dtype)
    c = np.empty((1), dtype)
    cext.bipartite_graph_calculate_hamiltonian(cobj, h0, h1, J, c, b0, b1, W, dtype)
    return h0, h1, J, c[0]


def bipartite_graph_calculate_E_from_spin(cext, cobj, h0, h1, J, c, q0, q1, dtype):
    h0, h1, J = common.fix_type([h0, h1, J], dtype)
    _check_bipartite(h0, h1, J)
    E = np.ndarray((1), dtype)
    cext.bipartite_graph_calculate_E_from_spin(cobj, E, h0, h1, J, dtype(c), _as_i8(q0), _as_i8(q1), dtype)
    return E[0]


def bipartite_graph_batch_calculate_E_from_spin(cext, cobj, h0, h1, J, c, q0, q1, dtype):
    h0, h1, J = common.fix_type([h0, h1, J], dtype)
    _check_bipartite(h0, h1, J)
    q0, q1 = _batch(q0), _batch(q1)
    if q0.shape[0] != q1.shape[0]:
        raise RuntimeError("Different batch dims between q0 and q1.")
    E = np.empty((q0.shape[0]), dtype)
    cext.bipartite_graph_batch_calculate_E_from_spin(cobj, E, h0, h1, J, dtype(c), q0, q1, dtype)
    return E
