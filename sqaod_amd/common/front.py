"""The Python front ends of the rocm solvers, assembled from one table.

Every sqaod solver object is a thin layer over its device extension: it
checks and converts the caller's arguments, then calls the extension with
(handle, arguments..., dtype) and hands the result back
(sqaodpy/sqaod/common/dense_graph_annealer_base.py:7-101,
bipartite_graph_annealer_base.py:8-111, dense_graph_bf_searcher_base.py:6-70,
bipartite_graph_bf_searcher_base.py:7-80 hold the reference's four).  Here
that layer is written once: `Front` owns the handle, an `Op` describes one
method (the extension function it calls, how its arguments are adapted and
how its result is shaped), and `assemble` builds a solver class from a list
of Ops.  The method names and argument meanings are the reference's, since
they are the interface callers program against.
"""
import numpy as np

from . import common
from . import preference as pref


class Front:
    """One solver handle (self._cobj) of an extension namespace (self._cext).
    Subclasses create the handle, then call Front.__init__ with the problem
    arrays (or None) and the preferences."""

    def __init__(self, cext, dtype, problem, optimize, prefdict):
        self._cext = cext
        self.dtype = dtype
        self._optimize = optimize
        if problem is not None and problem[-1] is not None:
            self.set_qubo(*problem, optimize)
        self.set_preferences(prefdict)

    def __del__(self):
        cobj = getattr(self, "_cobj", None)
        if cobj is not None:
            self._cobj = None
            self._cext.delete(cobj, self.dtype)

    def get_optimize_dir(self):
        return self._optimize

    def set_preferences(self, prefdict=None, **prefs):
        merged = dict(prefdict or {})
        merged.update(prefs)
        self._cext.set_preferences(self._cobj, merged, self.dtype)


class Op:
    """Method `name` of a solver: calls extension function `ext` (default: the
    same name) with the handle, the arguments as `adapt(solver, *args)`
    returns them (default: unchanged), and the dtype; `shape(solver, r)`
    post-processes the result."""

    def __init__(self, name, ext=None, adapt=None, shape=None, doc=None):
        self.name, self.ext, self.adapt, self.shape, self.doc = name, ext or name, adapt, shape, doc

    def method(self):
        ext, adapt, shape = self.ext, self.adapt, self.shape

        def call(solver, *args, **kwargs):
            if adapt is not None:
                args = adapt(solver, *args, **kwargs)
            elif kwargs:
                raise TypeError(f"{ext}() takes no keyword arguments")
            result = getattr(solver._cext, ext)(solver._cobj, *args, solver.dtype)
            return shape(solver, result) if shape is not None else result

        call.__name__ = self.name
        call.__doc__ = self.doc
        return call


def assemble(name, ops, doc, extra=None):
    """A Front subclass with one method per Op (and any `extra` attributes)."""
    ns = {op.name: op.method() for op in ops}
    ns.update(extra or {})
    ns["__doc__"] = doc
    return type(name, (Front,), ns)


# ---------------------------------------------------------------- argument adapters

def _square_symmetric(W, dtype):
    W = np.asarray(W)
    if W.ndim != 2 or W.shape[0] != W.shape[1]:
        raise RuntimeError("W must be a square matrix.")
    return common.symmetrize(common.fix_type(W, dtype))


def _dense_qubo(solver, W, optimize=pref.minimize):
    W = _square_symmetric(W, solver.dtype)
    solver._optimize = optimize
    return W, optimize


def _bipartite_qubo(solver, b0, b1, W, optimize=pref.minimize):
    b0, b1, W = (np.asarray(v) for v in (b0, b1, W))
    if b0.ndim != 1 or b1.ndim != 1 or W.shape != (b1.shape[0], b0.shape[0]):
        raise RuntimeError("Shape mismatch: W must be N1 x N0 for b0[N0], b1[N1].")
    solver._optimize = optimize
    return (*(np.ascontiguousarray(v, dtype=solver.dtype) for v in (b0, b1, W)), optimize)


def _dense_hamiltonian(solver, h, J, c):
    solver._optimize = pref.minimize
    h = np.ascontiguousarray(h, dtype=solver.dtype)
    return h, _square_symmetric(J, solver.dtype), solver.dtype(c)


def _bipartite_hamiltonian(solver, h0, h1, J, c):
    solver._optimize = pref.minimize
    h0, h1, J = (np.ascontiguousarray(v, dtype=solver.dtype) for v in (h0, h1, J))
    return h0, h1, J, solver.dtype(c)


def _int8(v):
    return np.ascontiguousarray(v, dtype=np.int8)


def _spins(solver, q):
    return (_int8(q),)


def _spin_rows(solver, qset):
    return ([_int8(q) for q in qset],)


def _spin_pair(solver, qpair):
    return ((_int8(qpair[0]), _int8(qpair[1])),)


def _spin_pairs(solver, qpairs):
    return ([(_int8(a), _int8(b)) for a, b in qpairs],)


def _in_dtype(solver, *scalars):
    """G, beta (annealOneStep, getSystemE) passed in the solver's real type."""
    return tuple(solver.dtype(v) for v in scalars)


def _dense_hamiltonian_out(solver):
    """get_hamiltonian(): fresh (h, J, c) buffers the extension fills."""
    N = solver.get_problem_size()
    out = (np.empty(N, solver.dtype), np.empty((N, N), solver.dtype), np.empty(1, solver.dtype))
    solver._cext.get_hamiltonian(solver._cobj, *out, solver.dtype)
    return out[0], out[1], out[2][0]


def _bipartite_hamiltonian_out(solver):
    N0, N1 = solver.get_problem_size()
    out = (np.empty(N0, solver.dtype), np.empty(N1, solver.dtype), np.empty((N1, N0), solver.dtype),
           np.empty(1, solver.dtype))
    solver._cext.get_hamiltonian(solver._cobj, *out, solver.dtype)
    return out[0], out[1], out[2], out[3][0]


def _search_all(solver):
    """search(): prepare, scan range after range until done, make_solution
    (DenseGraphBFSearcher::search, sqaodc/common/Solver.cpp:214-218)."""
    solver.prepare()
    while not solver.search_range()[0]:
        pass
    solver.make_solution()


def _search_in_library(solver):
    """The same search as one library call."""
    solver._cext.search(solver._cobj, solver.dtype)


# ---------------------------------------------------------------- the solvers

_COMMON = [Op("get_problem_size"), Op("get_preferences"), Op("get_E"), Op("get_x"), Op("prepare"),
           Op("calculate_E"), Op("make_solution")]

_ANNEALING = [Op("seed"), Op("randomize_spin"), Op("get_q"),
              Op("anneal_one_step", adapt=_in_dtype, doc="One sweep at transverse field G and inverse temperature beta."),
              Op("get_system_E", adapt=_in_dtype, doc="Mean energy of the Suzuki-Trotter system at (G, beta).")]

DenseGraphAnnealerBase = assemble(
    "DenseGraphAnnealerBase",
    _COMMON + _ANNEALING + [Op("set_qubo", adapt=_dense_qubo), Op("set_hamiltonian", adapt=_dense_hamiltonian),
                            Op("set_q", adapt=_spins), Op("set_qset", adapt=_spin_rows)],
    "Dense-graph annealer front end: W is a symmetric N x N QUBO.",
    {"get_hamiltonian": _dense_hamiltonian_out})

BipartiteGraphAnnealerBase = assemble(
    "BipartiteGraphAnnealerBase",
    _COMMON + _ANNEALING + [Op("set_qubo", adapt=_bipartite_qubo),
                            Op("set_hamiltonian", adapt=_bipartite_hamiltonian),
                            Op("set_q", adapt=_spin_pair), Op("set_qset", adapt=_spin_pairs)],
    "Bipartite-graph annealer front end: b0[N0], b1[N1], W N1 x N0; spins and x come in (x0, x1) pairs.",
    {"get_hamiltonian": _bipartite_hamiltonian_out})

DenseGraphBFSearcherBase = assemble(
    "DenseGraphBFSearcherBase",
    _COMMON + [Op("set_qubo", adapt=_dense_qubo),
               Op("search_range", doc="Scans the next span of x; returns (done, end of the span).")],
    "Dense-graph brute-force searcher front end.",
    {"search": _search_all, "_search": _search_in_library})

BipartiteGraphBFSearcherBase = assemble(
    "BipartiteGraphBFSearcherBase",
    _COMMON + [Op("set_qubo", adapt=_bipartite_qubo),
               Op("search_range", doc="Scans the next rectangle of pairs; returns (done, x0 reached, x1 reached).")],
    "Bipartite-graph brute-force searcher front end; get_x gives (x0, x1) pairs.",
    {"search": _search_all, "_search": _search_in_library})
