"""CPU checks of the fp8 form's numpy restatement (tests/fp8_ref.py), so the
GPU tests that compare the device's quantized J' against it rest on a pinned
definition: element rounding equals torch's float8_e4m3fn conversion, J' is
symmetric and stored exactly in both rows' grids, the row scale is the
smallest that fits, and quarter-integer J is kept exactly."""
import numpy as np
import pytest
import torch

import fp8_ref
from problems import dense_int_W, dense_normal_W, SEED_BASE


def test_e4m3_rounding_matches_torch():
    rng = np.random.default_rng(SEED_BASE)
    v = np.concatenate([
        rng.uniform(-448, 448, 20000),
        rng.standard_normal(20000) * 2.0 ** rng.integers(-12, 8, 20000),
        # every e4m3 value and every midpoint between neighbours (ties)
        np.ldexp(np.arange(0, 8), -9), np.ldexp(np.arange(0, 8) + 0.5, -9),
        np.ldexp(1 + np.arange(0, 8) / 8, 0), np.ldexp(1 + (np.arange(0, 8) + 0.5) / 8, 3),
    ]).astype(np.float32)
    v = v[np.abs(v) <= 448]  # the domain: callers scale into range (torch gives NaN past it)
    ref = torch.from_numpy(v).to(torch.float8_e4m3fn).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(fp8_ref.e4m3_round(v.astype(np.float64)), ref)


def test_e4m3_value_set():
    """The 126 positive finite e4m3 values are exactly the fixed points."""
    codes = np.arange(0, 127, dtype=np.uint8)  # 0x7F is NaN
    vals = torch.from_numpy(codes).view(torch.float8_e4m3fn).to(torch.float32).numpy().astype(np.float64)
    assert vals.max() == 448.0
    assert np.all(fp8_ref.is_e4m3(vals)) and np.all(fp8_ref.is_e4m3(-vals))
    assert not fp8_ref.is_e4m3(np.array([1.0625, 2.0 ** -10, 480.0])).any()


@pytest.mark.parametrize("N", [64, 300])
def test_quantized_J_is_symmetric_and_exact_in_both_rows(N):
    W = dense_normal_W(N, SEED_BASE + N)
    J = -0.25 * W.astype(np.float64)
    # rows of very different magnitudes, so max(a_i, a_j) matters
    s = np.ldexp(1.0, np.arange(N) % 7 - 3)
    J = J * s[:, None] * s[None, :]
    Jp, a = fp8_ref.quantize(J)
    assert np.array_equal(Jp, Jp.T)
    for i in range(0, N, 7):
        assert np.all(fp8_ref.is_e4m3(np.ldexp(Jp[i], -a[i]))), i
    # smallest exponent that fits each row
    rmax = np.abs(J).max(axis=1)
    assert np.all(rmax <= 448 * np.ldexp(1.0, a))
    assert np.all(rmax > 448 * np.ldexp(1.0, a - 1))
    # error at most half a quantum of the coarser row grid (4 significant bits)
    aij = np.maximum(a[:, None], a[None, :])
    tol = np.maximum(np.abs(J) * 2.0 ** -4, np.ldexp(1.0, aij - 10))
    assert np.all(np.abs(Jp - J) <= tol)


def test_quarter_integer_J_is_kept_exactly():
    W = dense_int_W(256, SEED_BASE + 3).astype(np.float64)
    J = -0.25 * W
    Jp, _ = fp8_ref.quantize(J)
    assert np.array_equal(Jp, J)


def test_exponent_floor_bounds_the_spread():
    J = np.diag([1000.0, 1e-9, 0.0, 2.0])
    a = fp8_ref.row_exponents(J)
    assert a[0] == 2 and a[1] == a[2] == a[0] - fp8_ref.EXP_SPREAD and a[3] == -7


def _host_conversions(tmp_path_factory):
    """The e4m3 conversions of sqaod_amd/csrc/fp8.h (the ones the device
    quantizer runs), compiled for the host."""
    import ctypes
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    d = tmp_path_factory.mktemp("fp8conv")
    src = d / "conv.cpp"
    src.write_text(
        '#include <cstdint>\n#define __host__\n#define __device__\n'
        f'#include "{root}/sqaod_amd/csrc/fp8.h"\n'
        'extern "C" void enc(const float *v, uint8_t *o, long n) { for (long i = 0; i < n; ++i) '
        'o[i] = sqaod_rocm::floatToFp8(v[i]); }\n'
        'extern "C" void dec(const uint8_t *b, float *o, long n) { for (long i = 0; i < n; ++i) '
        'o[i] = sqaod_rocm::fp8ToFloat(b[i]); }\n')
    so = d / "conv.so"
    subprocess.run(["g++", "-O2", "-std=c++20", "-shared", "-fPIC", str(src), "-o", str(so)], check=True)
    return ctypes.CDLL(str(so))


@pytest.fixture(scope="module")
def conv(tmp_path_factory):
    return _host_conversions(tmp_path_factory)


def test_device_conversions_match_torch(conv):
    import ctypes
    rng = np.random.default_rng(SEED_BASE + 1)
    v = np.concatenate([rng.uniform(-448, 448, 200000),
                        rng.standard_normal(200000) * 2.0 ** rng.integers(-14, 8, 200000),
                        np.ldexp(np.arange(0, 16) / 2, -9)]).astype(np.float32)
    v = v[np.abs(v) <= 448]
    out = np.empty(v.size, np.uint8)
    conv.enc(v.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), ctypes.c_long(v.size))
    ref = torch.from_numpy(v).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    # -0.0 and +0.0 encode differently; both decode to zero
    diff = out != ref
    assert np.all((v[diff] == 0) | (np.abs(v[diff]) < 2.0 ** -10)), v[diff][:5]
    codes = np.array([c for c in range(256) if c & 0x7F != 0x7F], np.uint8)
    dec = np.empty(codes.size, np.float32)
    conv.dec(codes.ctypes.data_as(ctypes.c_void_p), dec.ctypes.data_as(ctypes.c_void_p), ctypes.c_long(codes.size))
    assert np.array_equal(dec, torch.from_numpy(codes).view(torch.float8_e4m3fn).to(torch.float32).numpy())
