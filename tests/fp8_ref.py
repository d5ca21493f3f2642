"""numpy restatement of the fp8 (e4m3) precision form of the dense annealer's
J, used by the tests to check the device's quantization.  The form has no
counterpart in the reference (sqaod has fp32/fp64 only); its definition:

  * e4m3 = OCP FP8 E4M3 ("e4m3fn"): 4 significant bits, normal exponents
    -6..8, subnormal quantum 2^-9, largest value 448; rounding is to nearest,
    ties to even.
  * row i of J gets the scale 2^a_i with a_i the smallest integer such that
    max_j |J[i][j]| <= 448 * 2^a_i; all a_i are then raised to at least
    (largest a) - 15, and all-zero rows take that floor.
  * J'[i][j] = J[i][j] rounded on the e4m3 grid scaled by 2^max(a_i, a_j),
    so J' is symmetric whenever J is, and J'[i][j] / 2^a_i is an exact e4m3
    value (the byte row i stores).
"""
import numpy as np

E4M3_MAX = 448.0
EXP_SPREAD = 15


def e4m3_round(v):
    """Nearest-even e4m3 value of each element of v (float64, |v| <= 448)."""
    v = np.asarray(v, np.float64)
    a = np.abs(v)
    _, e = np.frexp(np.where(a > 0, a, 1.0))  # a = f 2^e, f in [0.5, 1): binade 2^(e-1)
    quantum = np.ldexp(1.0, np.maximum(e - 1, -6) - 3)
    r = np.round(a / quantum) * quantum  # np.round: half to even
    return np.sign(v) * np.minimum(r, E4M3_MAX)


def is_e4m3(v):
    v = np.asarray(v, np.float64)
    return (np.abs(v) <= E4M3_MAX) & (e4m3_round(v) == v)


def row_exponents(J):
    """a_i of the definition above (int64 array)."""
    J = np.asarray(J, np.float64)
    rmax = np.max(np.abs(J), axis=1)
    f, e = np.frexp(np.where(rmax > 0, rmax, 1.0))
    a = np.where(f <= 0.875, e - 9, e - 8)
    nz = rmax > 0
    amax = int(a[nz].max()) if nz.any() else 0
    a = np.where(nz, a, amax - EXP_SPREAD)
    return np.maximum(a, amax - EXP_SPREAD).astype(np.int64)


def quantize(J):
    """(J', a) for a symmetric J."""
    J = np.asarray(J, np.float64)
    a = row_exponents(J)
    aij = np.maximum(a[:, None], a[None, :])
    Jp = np.ldexp(e4m3_round(np.ldexp(J, -aij)), aij)
    return Jp, a
