"""Many dense QUBO instances annealed together on the rocm device.

There is no counterpart in sqaod: a DenseBatch holds float32 dense annealers
(one per instance, any N and m) and advances them through one schedule in
shared launches -- a grouped field GEMM, grouped Philox streams and one
persistent sweep over every (instance, slice) -- optionally captured as one
hipGraph per schedule.  Every instance ends exactly where a
dense_graph_annealer with the same W, m, seed and preferences ends after the
same anneal_one_step calls.
"""
import numpy as np

from ._cext import rocm_dense_batch as cext


class DenseBatch:
    def __init__(self, **prefs):
        self._cobj = cext.new()
        if prefs:
            cext.set_preferences(self._cobj, prefs)

    def __del__(self):
        if getattr(self, "_cobj", None) is not None:
            cext.delete(self._cobj)
            self._cobj = None

    def set_preferences(self, prefs=None, **kw):
        p = dict(prefs or {})
        p.update(kw)
        cext.set_preferences(self._cobj, p)

    def add_qubo(self, W, n_trotters, seed, maximize=False):
        """Adds an instance (W symmetric, N x N); returns its index."""
        return cext.add_qubo(self._cobj, W, n_trotters, seed, maximize)

    def __len__(self):
        return cext.size(self._cobj)

    def prepare(self):
        """prepare() and randomize_spin() of every instance."""
        cext.prepare(self._cobj)

    def anneal(self, G, beta, graph=True):
        """Steps k = 0..len(G)-1 with (G[k], beta[k]) on every instance;
        returns the schedule's device time in ms."""
        G = np.atleast_1d(np.asarray(G, np.float64))
        beta = np.broadcast_to(np.asarray(beta, np.float64), G.shape)
        return cext.anneal(self._cobj, G, beta, graph)

    def anneal_one_step(self, G, beta):
        return self.anneal([G], [beta], graph=False)

    def autotune(self, G, beta, steps=2):
        """Times candidate launch shapes and keeps the fastest
        (waves per block, largest-first order); results are unaffected."""
        return cext.autotune(self._cobj, G, beta, steps)

    def attempts_per_step(self):
        return cext.attempts_per_step(self._cobj)

    def shape(self, i):
        return cext.shape(self._cobj, i)

    def get_E(self, i):
        return cext.get_E(self._cobj, i)

    def get_x(self, i):
        return cext.get_x(self._cobj, i)

    def get_q(self, i):
        return cext.get_q(self._cobj, i)


def dense_batch(**prefs):
    return DenseBatch(**prefs)
