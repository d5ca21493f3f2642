"""rocm dense graph brute-force searcher
(sqaodpy/sqaod/cuda/dense_graph_bf_searcher.py)."""
import numpy as np

from ..common.front import DenseGraphBFSearcherBase
from ..common.preference import minimize
from ._cext import rocm_dg_bf_searcher as cext
from . import device


class DenseGraphBFSearcher(DenseGraphBFSearcherBase):
    """Dense graph brute-force searcher on the rocm device."""

    def __init__(self, W, optimize, dtype, prefdict):
        self._cobj = cext.new(dtype)
        dev = device.get_active_device()
        cext.assign_device(self._cobj, dev._cobj, dtype)
        self._device = dev
        DenseGraphBFSearcherBase.__init__(self, cext, dtype, (W,), optimize, prefdict)


def dense_graph_bf_searcher(W=None, optimize=minimize, dtype=np.float64, **prefs):
    """Factory of sqaod_amd.rocm.DenseGraphBFSearcher (same signature as
    sqaod.cuda.dense_graph_bf_searcher)."""
    return DenseGraphBFSearcher(W, optimize, dtype, prefs)
