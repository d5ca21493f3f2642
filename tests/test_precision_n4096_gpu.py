"""The precision configuration at its full size: dense SQA with N = 4096,
m = 256, W symmetric i.i.d. N(0, 1) float32, the whole schedule (G = 5.0 *
0.97^k while G > 0.01, 205 steps, beta = 50), in each field form of the
rocm dense annealer (fp32, bf16 hi + lo, fp8 e4m3 with a power-of-two scale
per row).

For every form:
* get_E is the FP64 recompute of the returned spins: it equals a host FP64
  x^T W x of the same x (the reference's energy, SharedFormulas.cpp:70-83)
  within 1e-5 relative (fp32 rounding of the result is 6e-8);
* get_form_energy_error equals the host's max_y |E_J'(q_y) - E_J(q_y)|, J'
  being the J the form anneals (bf16: float(hi) + float(lo), hi = bf16(J),
  lo = bf16(J - hi); fp8: the e4m3 J' of tests/fp8_ref.py);
* the form's stated bound holds: that error is at most 1e-4 of |E| for fp32
  and bf16x2, at most 1e-3 of |E| for fp8 (e4m3 keeps 4 significant bits, so
  J' differs from J by up to 2^-4 relative per element; 4.9e-4 measured).
"""
import numpy as np
import pytest

import fp8_ref
from problems import dense_normal_W, schedule

pytestmark = pytest.mark.gpu

N, M, BETA, SEED = 4096, 256, 50.0, 1133557
BOUND = {"fp32": 1e-4, "bf16x2": 1e-4, "fp8": 1e-3}


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


@pytest.fixture(scope="module")
def problem():
    W = dense_normal_W(N, SEED)
    J = (-0.25 * W).astype(np.float32)  # SharedFormulas.cpp:36-55, exact in float
    np.fill_diagonal(J, 0.0)
    return W, J


def bf16_rne(a):
    """float32 -> bfloat16 (round to nearest even) -> float32, finite inputs."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def form_J(J, form):
    if form == "fp32":
        return J.astype(np.float64)
    if form == "bf16x2":
        hi = bf16_rne(J)
        lo = bf16_rne((J - hi).astype(np.float32))
        return (hi.astype(np.float32) + lo.astype(np.float32)).astype(np.float64)
    Jp, _ = fp8_ref.quantize(J.astype(np.float64))
    return Jp


@pytest.mark.parametrize("form", ["fp32", "bf16x2", "fp8"])
def test_full_schedule_energy_precision(rocm, problem, form):
    W, J = problem
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=M, field_precision=form)
    an.seed(1000)
    an.prepare()
    an.randomize_spin()
    for G in schedule():
        an.anneal_one_step(G, BETA)
    an.make_solution()
    E = an.get_E().astype(np.float64)
    x = np.stack(an.get_x()).astype(np.float64)
    q = 2.0 * x - 1.0
    E_host = np.einsum("yi,yi->y", x @ W.astype(np.float64), x)
    rel = np.max(np.abs(E - E_host) / np.abs(E_host))
    assert rel <= 1e-5, f"{form}: get_E vs host FP64 {rel:.3g}"
    # the form's own energy error against the J as given
    D = form_J(J, form) - J.astype(np.float64)
    host_err = float(np.max(np.abs(np.einsum("yi,yi->y", q @ D, q))))
    dev_err = an.get_form_energy_error()
    assert dev_err == pytest.approx(host_err, rel=1e-6, abs=1e-6), form
    bound = BOUND[form] * float(np.max(np.abs(E_host)))
    assert host_err <= bound, f"{form}: form energy error {host_err:.4g} over the stated bound {bound:.4g}"
    # the anneal actually progressed: far below the random-spin energy (~0)
    assert E_host.max() < -1.4e5
