"""Array helpers the solver front ends and callers share; the names are those
of sqaodpy/sqaod/common/common.py, which callers import."""
import numbers

import numpy as np

__all__ = ["is_symmetric", "is_triangular", "symmetrize", "generate_random_symmetric_W", "fix_type",
           "create_bitset_sequence", "bit_to_spin", "spin_to_bit", "clone_as_ndarray"]

# (rtol, atol) by precision: how far W may be from W^T and still count as symmetric
_TOLERANCE = {np.dtype(np.float64): (1.0e-5, 1.0e-8)}
_TOLERANCE_DEFAULT = (1.0e-4, 1.0e-5)


def _tolerance(mat):
    return _TOLERANCE.get(np.asarray(mat).dtype, _TOLERANCE_DEFAULT)


def is_symmetric(mat):
    rtol, atol = _tolerance(mat)
    return np.allclose(mat, np.transpose(mat), rtol, atol)


def is_triangular(mat):
    """One strict triangle is zero (within the precision's rtol)."""
    tol = _tolerance(mat)[0]
    mag = np.abs(np.asarray(mat))
    return bool(np.all(np.triu(mag, 1) <= tol) or np.all(np.tril(mag, -1) <= tol))


def symmetrize(mat):
    """A symmetric W passes unchanged; a W that is symmetric only within the
    precision's tolerance is made exactly symmetric from its upper triangle
    (the rocm kernels rely on W == W^T); a triangular W becomes (W + W^T)/2;
    anything else is an error (sqaodpy/sqaod/common/common.py:52-59 accepts
    the same three cases)."""
    if np.array_equal(mat, np.transpose(mat)):
        return mat
    if is_symmetric(mat):
        upper = np.triu(mat)
        return upper + np.triu(mat, 1).T
    if is_triangular(mat):
        return (mat + np.transpose(mat)) * 0.5
    raise RuntimeError("given matrix is not triangular nor symmetric.")


def generate_random_symmetric_W(N, wmin=-0.5, wmax=0.5, dtype=np.float64):
    """A symmetric N x N matrix with entries uniform in [wmin, wmax), from
    numpy's global generator."""
    U = np.random.random_sample((N, N))
    S = np.triu(U) + np.triu(U, 1).T
    return np.asarray(wmin + (wmax - wmin) * S, dtype)


def clone_as_ndarray(var, dtype):
    return np.array([var] if isinstance(var, numbers.Number) else var, dtype)


def _as_array(var, dtype):
    if isinstance(var, np.ndarray) and var.dtype == dtype and var.flags["C_CONTIGUOUS"]:
        return var
    return np.ascontiguousarray(clone_as_ndarray(var, dtype))


def fix_type(var, dtype):
    """var as a C-contiguous ndarray of dtype, or, for a list / tuple, a list
    of its elements each converted so; arrays that already are one are
    returned as they are."""
    if isinstance(var, (list, tuple)):
        return [_as_array(v, dtype) for v in var]
    return _as_array(var, dtype)


def create_bitset_sequence(vals, nbits):
    """Packed x values -> rows of int8 bits, element pos = bit nbits-1-pos
    (sqaodc/common/Common.cpp:79-93)."""
    packed = np.atleast_1d(np.asarray(vals, dtype=np.uint64))
    weights = np.uint64(1) << np.arange(nbits - 1, -1, -1, dtype=np.uint64)
    return ((packed[:, None] & weights[None, :]) != 0).astype(np.int8)


def bit_to_spin(bits):
    """x in {0, 1} -> q = 2x - 1 in {-1, +1}."""
    return (2 * np.asarray(bits, np.int8) - 1).astype(np.int8)


def spin_to_bit(spin):
    """q in {-1, +1} -> x = (q + 1) / 2 in {0, 1}."""
    return (np.asarray(spin, np.int8) > 0).astype(np.int8)
