"""wildqat's optimizer driving the rocm dense annealer
(the counterpart of sqaodpy/sqaod/wildqat)."""
from .opt import opt  # noqa: F401
