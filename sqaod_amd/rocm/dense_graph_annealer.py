"""rocm dense graph annealer (sqaodpy/sqaod/cuda/dense_graph_annealer.py:1-34)."""
import numpy as np

from ..common.front import DenseGraphAnnealerBase
from ..common.preference import minimize
from ._cext import rocm_dg_annealer as cext
from . import device


class DenseGraphAnnealer(DenseGraphAnnealerBase):
    """Dense graph annealer on the rocm device."""

    def __init__(self, W, optimize, dtype, prefdict):
        self._cobj = cext.new(dtype)
        dev = device.get_active_device()
        cext.assign_device(self._cobj, dev._cobj, dtype)
        self._device = dev
        DenseGraphAnnealerBase.__init__(self, cext, dtype, (W,), optimize, prefdict)

    def get_counters(self):
        """Accepted flips, committed batches and blocked waits since prepare()."""
        return cext.get_counters(self._cobj, self.dtype)

    def get_phase_times(self):
        """With preference profile=1: {phase: (device ms, launches)} for the
        refresh, rng and sweep phases since the last call."""
        return cext.get_phase_times(self._cobj, self.dtype)

    def get_fields(self):
        """Local fields F = J q_y (m x N) of the local-field sweep."""
        return cext.get_fields(self._cobj, self.dtype)

    def trace_system_E(self, capacity):
        """From now on every anneal_one_step(G, beta) records get_system_E(G,
        beta) of the state after the step, on the device and without a host
        synchronisation (the first `capacity` steps); 0 turns it off."""
        cext.set_system_E_trace(self._cobj, capacity, self.dtype)
        self._trace_cap = int(capacity)

    def get_system_E_trace(self):
        """The recorded values (float64 array) and the number of steps traced."""
        return cext.get_system_E_trace(self._cobj, getattr(self, "_trace_cap", 0), self.dtype)

    def get_E_fp64(self):
        """E of every slice from the last energy evaluation, in float64 before
        its rounding to dtype (exact on integer problems)."""
        return cext.get_E_fp64(self._cobj, self.dtype)

    def get_energy_rounding_error(self):
        """max |E(dtype) - E(fp64)| of the last energy evaluation."""
        return cext.get_energy_rounding_error(self._cobj, self.dtype)

    def get_form_energy_error(self):
        """max over slices of |E_form(q) - E(q)| in FP64, E_form with the J the
        field precision form anneals (bf16 hi + lo, fp8 J'), for the current q."""
        return cext.get_form_energy_error(self._cobj, self.dtype)


def dense_graph_annealer(W=None, optimize=minimize, dtype=np.float64, **prefs):
    """Factory of sqaod_amd.rocm.DenseGraphAnnealer (same signature as
    sqaod.cuda.dense_graph_annealer)."""
    return DenseGraphAnnealer(W, optimize, dtype, prefs)
