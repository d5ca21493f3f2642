"""rocm formulas (sqaodpy/sqaod/cuda/formulas.py): module-level functions on
the active device, backed by one formulas object created at import."""
import atexit
import sys

from ..common import formulas_base as base
from ._cext import rocm_formulas as cext
from . import device


def dense_graph_calculate_E(W, x, dtype):
    return base.dense_graph_calculate_E(cext, _obj(), W, x, dtype)


def dense_graph_batch_calculate_E(W, x, dtype):
    return base.dense_graph_batch_calculate_E(cext, _obj(), W, x, dtype)


def dense_graph_calculate_hamiltonian(W, dtype):
    return base.dense_graph_calculate_hamiltonian(cext, _obj(), W, dtype)


def dense_graph_calculate_E_from_spin(h, J, c, q, dtype):
    return base.dense_graph_calculate_E_from_spin(cext, _obj(), h, J, c, q, dtype)


def dense_graph_batch_calculate_E_from_spin(h, J, c, q, dtype):
    return base.dense_graph_batch_calculate_E_from_spin(cext, _obj(), h, J, c, q, dtype)


def bipartite_graph_calculate_E(b0, b1, W, x0, x1, dtype):
    return base.bipartite_graph_calculate_E(cext, _obj(), b0, b1, W, x0, x1, dtype)


def bipartite_graph_batch_calculate_E(b0, b1, W, x0, x1, dtype):
    return base.bipartite_graph_batch_calculate_E(cext, _obj(), b0, b1, W, x0, x1, dtype)


def bipartite_graph_batch_calculate_E_2d(b0, b1, W, x0, x1, dtype):
    return base.bipartite_graph_batch_calculate_E_2d(cext, _obj(), b0, b1, W, x0, x1, dtype)


def bipartite_graph_calculate_hamiltonian(b0, b1, W, dtype):
    return base.bipartite_graph_calculate_hamiltonian(cext, _obj(), b0, b1, W, dtype)


def bipartite_graph_calculate_E_from_spin(h0, h1, J, c, q0, q1, dtype):
    return base.bipartite_graph_calculate_E_from_spin(cext, _obj(), h0, h1, J, c, q0, q1, dtype)


def bipartite_graph_batch_calculate_E_from_spin(h0, h1, J, c, q0, q1, dtype):
    return base.bipartite_graph_batch_calculate_E_from_spin(cext, _obj(), h0, h1, J, c, q0, q1, dtype)


_this = sys.modules[__name__]
_this._cobj = None


def _obj():
    """The formulas object, created (and bound to the active device) on first use."""
    if _this._cobj is None:
        _this._cobj = cext.dg_formulas_new()
        cext.dg_formulas_assign_device(_this._cobj, device.get_active_device()._cobj)
    return _this._cobj


def unload():
    if _this._cobj is not None:
        cext.dg_formulas_delete(_this._cobj)
        _this._cobj = None


atexit.register(unload)
