"""Algorithm tags and optimize directions (sqaodpy/sqaod/common/preference.py:1-80)."""


class Algorithm:
    default = 'default'
    naive = 'naive'
    coloring = 'coloring'
    brute_force_search = 'brute_force_search'
    sa_default = 'sa_default'
    sa_naive = 'sa_naive'
    sa_coloring = 'sa_coloring'

    @staticmethod
    def is_sqa(algo):
        return algo in (Algorithm.default, Algorithm.naive, Algorithm.coloring)


algorithm = Algorithm()


class Minimize:
    """Tag telling solvers to minimize the QUBO energy."""

    @staticmethod
    def sign(v):
        return v.copy()

    @staticmethod
    def best(values):
        return min(values)

    @staticmethod
    def sort(values):
        return sorted(values)

    def __int__(self):
        return 0


class Maximize:
    """Tag telling solvers to maximize the QUBO energy."""

    @staticmethod
    def sign(v):
        return -v

    @staticmethod
    def best(values):
        return max(values)

    @staticmethod
    def sort(values):
        return sorted(values, reverse=True)

    def __int__(self):
        return 1


minimize = Minimize()
maximize = Maximize()
