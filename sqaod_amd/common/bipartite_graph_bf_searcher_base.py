"""Device-independent bipartite brute-force searcher front end
(sqaodpy/sqaod/common/bipartite_graph_bf_searcher_base.py:7-80)."""
from . import common
from . import preference as pref


class BipartiteGraphBFSearcherBase:

    def __init__(self, cext, dtype, b0, b1, W, optimize, prefdict):
        self._cext = cext
        self.dtype = dtype
        self._optimize = optimize
        if W is not None:
            self.set_qubo(b0, b1, W, optimize)
        self.set_preferences(prefdict)

    def __del__(self):
        if getattr(self, '_cobj', None) is not None:
            try:
                self._cext.delete(self._cobj, self.dtype)
            finally:
                self._cobj = None

    def set_qubo(self, b0, b1, W, optimize=pref.minimize):
        b0, b1, W = common.fix_type([b0, b1, W], self.dtype)
        self._dim = (b0.shape[0], b1.shape[0])
        self._cext.set_qubo(self._cobj, b0, b1, W, optimize, self.dtype)
        self._optimize = optimize

    def get_problem_size(self):
        return self._cext.get_problem_size(self._cobj, self.dtype)

    def set_preferences(self, prefdict=None, **prefs):
        if prefdict is not None:
            self._cext.set_preferences(self._cobj, prefdict, self.dtype)
        self._cext.set_preferences(self._cobj, prefs, self.dtype)

    def get_preferences(self):
        return self._cext.get_preferences(self._cobj, self.dtype)

    def get_optimize_dir(self):
        return self._optimize

    def get_E(self):
        return self._cext.get_E(self._cobj, self.dtype)

    def get_x(self):
        """List of (x0, x1) pairs."""
        return self._cext.get_x(self._cobj, self.dtype)

    def prepare(self):
        self._cext.prepare(self._cobj, self.dtype)

    def make_solution(self):
        self._cext.make_solution(self._cobj, self.dtype)

    def calculate_E(self):
        self._cext.calculate_E(self._cobj, self.dtype)

    def search_range(self):
        """Scans the next rectangle of pairs; returns (done, x0 reached, x1 reached)."""
        return self._cext.search_range(self._cobj, self.dtype)

    def search(self):
        self.prepare()
        while True:
            comp, curx0, curx1 = self._cext.search_range(self._cobj, self.dtype)
            if comp:
                break
        self.make_solution()

    def _search(self):
        # one call into the library
        self._cext.search(self._cobj, self.dtype)
