"""Seeded test problems shared by the CPU and GPU tests.

The dense weight law of the bit-exact configuration: symmetric W with
entries i.i.d. uniform integers in [-8, 8] (upper triangle incl. diagonal
drawn row-major from std::mt19937_64-equivalent numpy PCG64 seeded per
instance, then mirrored).  Integer W makes J = -W/4 quarter-integral, so every
field and energy is exact in float.
"""
import numpy as np

SEED_BASE = 1133557


def dense_int_W(N, seed, lo=-8, hi=8, dtype=np.float32):
    rng = np.random.default_rng(seed)
    U = rng.integers(lo, hi + 1, size=(N, N)).astype(np.float64)
    W = np.triu(U) + np.triu(U, 1).T
    return W.astype(dtype)


def dense_normal_W(N, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    U = rng.standard_normal((N, N))
    W = np.triu(U) + np.triu(U, 1).T
    return W.astype(dtype)


def schedule(G0=5.0, Gmin=0.01, tau=0.97):
    """G from 5.0 down to 0.01, multiplied by 0.97 per step (205 steps)."""
    out = []
    G = G0
    while G > Gmin:
        out.append(G)
        G *= tau
    return out


def dense_graph_8x8(dtype=np.float64):
    """sqaodpy/tests/example_problems.py:4-14: diagonal -32, off-diagonal 4."""
    W = np.full((8, 8), 4.0)
    np.fill_diagonal(W, -32.0)
    return W.astype(dtype)


def mat_symmetric_cycle(N, dtype=np.float64):
    """sqaodc/tests/utils.h:80-95 (testMatSymmetric): entries cycle through -2..2."""
    W = np.zeros((N, N))
    v = 0
    for i in range(N):
        for j in range(i, N):
            W[i, j] = W[j, i] = float(v % 5 - 2)
            v += 1
    return W.astype(dtype)


def quantize(W, dtype):
    """sqaodpy/tests/example_problems.py:16-18: multiples of 2^-14."""
    return np.asarray(np.rint(np.asarray(W, np.float64) * 16384) / 16384., dtype)


def dense_graph_random(N, seed, dtype=np.float64):
    """example_problems.py:20-22 with a seeded generator: U[-0.5, 0.5) symmetric, quantized."""
    rng = np.random.default_rng(seed)
    U = rng.random((N, N))
    W = np.triu(U) + np.triu(U, 1).T
    return quantize(W - 0.5, dtype)
