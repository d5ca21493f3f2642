"""rocm bipartite graph annealer (sqaodpy/sqaod/cuda/bipartite_graph_annealer.py)."""
import numpy as np

from ..common.front import BipartiteGraphAnnealerBase
from ..common.preference import minimize
from ._cext import rocm_bg_annealer as cext
from . import device


class BipartiteGraphAnnealer(BipartiteGraphAnnealerBase):
    """Bipartite graph annealer on the rocm device."""

    def __init__(self, b0, b1, W, optimize, dtype, prefdict):
        self._cobj = cext.new(dtype)
        dev = device.get_active_device()
        cext.assign_device(self._cobj, dev._cobj, dtype)
        self._device = dev
        BipartiteGraphAnnealerBase.__init__(self, cext, dtype, (b0, b1, W), optimize, prefdict)

    def get_E_fp64(self):
        """E of every slice from the last energy evaluation, in float64 before
        its rounding to dtype (exact on integer problems)."""
        return cext.get_E_fp64(self._cobj, self.dtype)


def bipartite_graph_annealer(b0=None, b1=None, W=None, optimize=minimize, dtype=np.float64, **prefs):
    """Factory of sqaod_amd.rocm.BipartiteGraphAnnealer (same signature as
    sqaod.cuda.bipartite_graph_annealer)."""
    return BipartiteGraphAnnealer(b0, b1, W, optimize, dtype, prefs)
