"""Serving anneal jobs with continuous batching on the rocm device.

There is no counterpart in sqaod.  A Server accepts dense and bipartite
jobs (float32 QUBOs, each with its own seed and G/beta schedule) at any
time; each tick() admits waiting jobs in arrival order while the active-job
limit and the device-memory budget allow, advances every active job by one
step together -- dense jobs in shared launches, bipartite jobs on parallel
streams -- and finishes the jobs whose schedule is done.  A finished job's
E and x equal those of the same job run alone (dense_graph_annealer /
bipartite_graph_annealer with the same seed and schedule).
"""
from ._cext import rocm_server as cext


class Server:
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

    def submit_dense(self, W, n_trotters, seed, G, beta):
        return cext.submit_dense(self._cobj, W, n_trotters, seed, G, beta)

    def submit_bipartite(self, b0, b1, W, n_trotters, seed, G, beta):
        return cext.submit_bipartite(self._cobj, b0, b1, W, n_trotters, seed, G, beta)

    def tick(self):
        """One step of every active job; returns the jobs active or pending."""
        return cext.tick(self._cobj)

    def status(self):
        return cext.status(self._cobj)

    def take_done(self):
        return cext.take_done(self._cobj)

    def result(self, job):
        """E, x (or x0, x1), times and attempts of a finished job; raises
        RuntimeError with the job's message if it failed."""
        return cext.result(self._cobj, job)

    def error(self, job):
        """None for a job that completed, else why it failed (admission, a
        step or its solution threw); the server went on with the others."""
        return cext.error(self._cobj, job)

    def release(self, job):
        cext.release(self._cobj, job)

    def now(self):
        """Seconds on the server's clock (the clock of the result times)."""
        return cext.now(self._cobj)

    def run_until_idle(self):
        while self.tick() > 0:
            pass


def server(**prefs):
    return Server(**prefs)
