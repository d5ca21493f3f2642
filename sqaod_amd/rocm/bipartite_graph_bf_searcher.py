"""rocm bipartite graph brute-force searcher
(sqaodpy/sqaod/cuda/bipartite_graph_bf_searcher.py)."""
import numpy as np

from ..common.front import BipartiteGraphBFSearcherBase
from ..common.preference import minimize
from ._cext import rocm_bg_bf_searcher as cext
from . import device


class BipartiteGraphBFSearcher(BipartiteGraphBFSearcherBase):
    """Bipartite graph brute-force searcher on the rocm device."""

    def __init__(self, b0, b1, W, optimize, dtype, prefdict):
        self._cobj = cext.new(dtype)
        dev = device.get_active_device()
        cext.assign_device(self._cobj, dev._cobj, dtype)
        self._device = dev
        BipartiteGraphBFSearcherBase.__init__(self, cext, dtype, (b0, b1, W), optimize, prefdict)


def bipartite_graph_bf_searcher(b0=None, b1=None, W=None, optimize=minimize, dtype=np.float64, **prefs):
    """Factory of sqaod_amd.rocm.BipartiteGraphBFSearcher (same signature as
    sqaod.cuda.bipartite_graph_bf_searcher)."""
    return BipartiteGraphBFSearcher(b0, b1, W, optimize, dtype, prefs)
