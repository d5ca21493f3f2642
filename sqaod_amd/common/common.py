"""Helpers shared by the solvers (sqaodpy/sqaod/common/common.py)."""
import numbers

import numpy as np

__all__ = ["is_symmetric", "is_triangular", "symmetrize", "generate_random_symmetric_W", "fix_type",
           "create_bitset_sequence", "bit_to_spin", "spin_to_bit", "clone_as_ndarray"]

rtol_fp64, atol_fp64 = 1.0e-5, 1.0e-8
rtol_fp32, atol_fp32 = 1.0e-4, 1.0e-5


def is_symmetric(mat):
    if mat.dtype == np.float64:
        rtol, atol = rtol_fp64, atol_fp64
    else:
        rtol, atol = rtol_fp32, atol_fp32
    return np.allclose(mat, mat.T, rtol, atol)


def is_triangular(mat):
    atol = rtol_fp64 if mat.dtype == np.float64 else rtol_fp32
    if not np.any(np.triu(mat, 1) > atol):
        return True
    return not np.any(np.tril(mat, -1) > atol)


def symmetrize(mat):
    """common.py:52-59: a (nearly) symmetric matrix passes; a triangular one
    is folded into (W + W^T)/2.  The rocm solvers need exact symmetry, so a
    matrix accepted as nearly symmetric is made exactly symmetric from its
    upper triangle."""
    if is_symmetric(mat):
        if np.array_equal(mat, mat.T):
            return mat
        return np.triu(mat) + np.triu(mat, 1).T
    if is_triangular(mat):
        return (mat + mat.T) * 0.5
    raise RuntimeError("given matrix is not triangular nor symmetric.")


def generate_random_symmetric_W(N, wmin=-0.5, wmax=0.5, dtype=np.float64):
    """common.py:67-74"""
    W = np.zeros((N, N), dtype)
    for i in range(0, N):
        for j in range(i, N):
            W[i, j] = np.random.random()
    W = W + np.tril(np.ones((N, N)), -1) * W.T
    W = W * (wmax - wmin) + wmin
    return W


def clone_as_ndarray(var, dtype):
    if isinstance(var, numbers.Number):
        return np.array([var], dtype)
    return np.array(var, dtype)


def fix_type(var, dtype):
    """Return var as a C-contiguous ndarray of dtype (common.py:79-108)."""
    if isinstance(var, (list, tuple)):
        return [fix_type(v, dtype) for v in var]
    if isinstance(var, np.ndarray) and var.dtype == dtype and var.flags['C_CONTIGUOUS']:
        return var
    return np.ascontiguousarray(clone_as_ndarray(var, dtype))


def create_bitset_sequence(vals, nbits):
    """Packed values -> int8 bit vectors, MSB first (common.py:124-135, Common.cpp:79-86)."""
    vals = np.asarray(vals, dtype=np.uint64)
    shifts = np.arange(nbits - 1, -1, -1, dtype=np.uint64)
    bits = (vals[:, None] >> shifts[None, :]) & np.uint64(1)
    return bits.astype(np.int8)


def bit_to_spin(bits):
    return np.asarray(bits, np.int8) * 2 - 1


def spin_to_bit(spin):
    return ((np.asarray(spin, np.int8) + 1) // 2).astype(np.int8)
