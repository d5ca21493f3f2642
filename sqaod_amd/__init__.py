"""sqaod_amd -- the rocm (MI355X, gfx950) device for sqaod's simulated quantum
annealing solvers.

Mirrors the top level of the reference Python package (sqaodpy/sqaod/__init__.py):
``algorithm``, ``minimize``/``maximize`` and the common helpers, plus the
device package ``sqaod_amd.rocm`` that plays the role of ``sqaod.cuda``.
"""
from .common.preference import algorithm, minimize, maximize  # noqa: F401
from .common.common import *  # noqa: F401,F403

__version__ = "1.0.0"


def is_rocm_available():
    """True when libsqaod_rocm.so is built and a GPU is visible."""
    try:
        from .rocm import _cext
        return _cext.device_count() > 0
    except Exception:
        return False
