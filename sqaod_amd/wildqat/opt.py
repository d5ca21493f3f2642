"""wildqat's `opt` on sqaod's dense annealer (sqaodpy/sqaod/wildqat/opt.py:1-67),
with the rocm device as the default package.

The reference adapter subclasses wildqat.opt, takes the QUBO and the
schedule parameters from it (qubo, Ts, Tf, R and ite for SA; Gs, Gf, tro,
Tf, R and ite for SQA), runs the annealer and records the system energy
after every step (SA) or after every G level (SQA) with get_system_E, which
on a device means a host round trip per step.  Here the annealer's system
energy trace records the same values on the device during the run and the
list is read once at the end; with a package whose annealer has no trace
the adapter calls get_system_E as the reference does.

wildqat itself is not a dependency: when it is importable, `opt` subclasses
wildqat.opt as the reference does; otherwise it holds the parameters the
adapter reads itself (with this module's defaults, listed in _Parameters).
"""
import numpy as np

from ..common import preference as pref

try:  # pragma: no cover - wildqat is optional
    import wildqat as _wildqat
    _Base = _wildqat.opt
except ImportError:
    _wildqat = None
    _Base = None


class _Parameters:
    """The attributes of wildqat.opt the adapter reads, for use without
    wildqat: the QUBO (a square nested list or array) and the schedules."""

    def __init__(self):
        self.qubo = []
        self.Ts = 5.0    # SA: initial temperature
        self.Tf = 0.02   # SA: final temperature; SQA: 1 / beta
        self.R = 0.95    # schedule ratio per level
        self.ite = 1000  # spin-flip attempts per level
        self.Gs = 10.0   # SQA: initial transverse field
        self.Gf = 0.02   # SQA: final transverse field
        self.tro = 8     # SQA: Trotter slices
        self.E = []

    def qi(self):
        """wildqat's hook after the QUBO is set; nothing to do without it."""
        return None


def _levels(start, stop, ratio):
    n, v = 0, start
    while v > stop:
        n += 1
        v *= ratio
    return n


class opt(_Base if _Base is not None else _Parameters):
    def __init__(self, pkg=None, dtype=np.float32):
        (_Base if _Base is not None else _Parameters).__init__(self)
        if pkg is None:
            from .. import rocm as pkg
        self.ann = pkg.dense_graph_annealer(dtype=dtype)

    def _has_qubo(self):
        return self.qubo is not None and len(self.qubo) != 0

    def _run(self, level_start, level_stop, n_per_level, step, record_each_step, beta_of):
        """Steps through the levels (level *= R while above level_stop),
        n_per_level steps each; returns the system energies the reference
        records: after every step or after the last step of every level."""
        n_levels = _levels(level_start, level_stop, self.R)
        traced = hasattr(self.ann, "trace_system_E")
        if traced:
            self.ann.trace_system_E(n_levels * n_per_level)
        E = []
        v = level_start
        while v > level_stop:
            for _ in range(n_per_level):
                step(v)
                if not traced and record_each_step:
                    E.append(self.ann.get_system_E(0., 0.))
            if not traced and not record_each_step:
                E.append(self.ann.get_system_E(v, beta_of(v)))
            v *= self.R
        if traced:
            values, _ = self.ann.get_system_E_trace()
            self.ann.trace_system_E(0)
            E = list(values) if record_each_step else list(values[n_per_level - 1::n_per_level])
        return E

    def sa(self):
        """SA on the QUBO (n_trotters = 1, sa_naive): T from Ts down to Tf,
        multiplied by R per level, ceil(ite / N) steps per level; self.E gets
        the energy after every step; returns the best slice's x."""
        if self._has_qubo():
            self.ann.set_qubo(np.array(self.qubo), pref.minimize)
            self.ann.set_preferences(algorithm=pref.algorithm.sa_naive, n_trotters=1)
        self.ann.prepare()
        self.ann.randomize_spin()
        N = self.ann.get_problem_size()
        n_iters = (self.ite + N - 1) // N
        self.E = self._run(self.Ts, self.Tf, n_iters, lambda T: self.ann.anneal_one_step(T, 1.), True,
                           lambda T: 1.)
        return self.ann.get_x()[0]

    def sqa(self):
        """SQA on the QUBO with tro Trotter slices: G from Gs down to Gf,
        multiplied by R per level, beta = 1 / Tf, ceil(ite / (N tro)) steps
        per level; self.E gets the system energy after every level; returns
        the x of every slice."""
        if self._has_qubo():
            self.ann.set_qubo(np.array(self.qubo), pref.minimize)
            self.ann.set_preferences(algorithm=pref.algorithm.default, n_trotters=self.tro)
            self.qi()
        self.ann.prepare()
        self.ann.randomize_spin()
        N = self.ann.get_problem_size()
        n_flips = N * self.tro
        n_iters = (self.ite + n_flips - 1) // n_flips
        beta = 1. / self.Tf
        self.E = self._run(self.Gs, self.Gf, n_iters, lambda G: self.ann.anneal_one_step(G, beta), False,
                           lambda G: beta)
        return self.ann.get_x()
