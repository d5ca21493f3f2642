"""The fp8 precision form of the dense annealer on the device: the e4m3
decode, the quantized J' (through the fields and the form energy error)
against the numpy restatement in fp8_ref.py, and trajectories that equal the
CPU annealer's bit for bit when J is exact in e4m3."""
import numpy as np
import pytest
import torch

import fp8_ref
from problems import dense_int_W, dense_normal_W, schedule, SEED_BASE

pytestmark = pytest.mark.gpu

BETA = 50.0


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def test_device_decode_table(rocm):
    """v_cvt_pk_f32_fp8 as the sweep uses it decodes every non-NaN e4m3 code
    exactly (OCP E4M3: bias 7, max 448)."""
    from sqaod_amd.rocm import _cext
    dev = _cext.fp8_decode_table()
    codes = np.array([c for c in range(256) if c & 0x7F != 0x7F], np.uint8)
    ref = torch.from_numpy(codes).view(torch.float8_e4m3fn).to(torch.float32).numpy()
    assert np.array_equal(dev[codes], ref)


def make(rocm, W, m, **prefs):
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m, **prefs)
    an.seed(23)
    an.prepare()
    an.randomize_spin()
    return an


@pytest.mark.parametrize("N,m", [(1024, 32), (700, 9)])
def test_fields_equal_host_product_with_quantized_J(oracle, rocm, N, m):
    """W ~ N(0,1): the fp8 fields equal q . J'^T (J' from fp8_ref) to fp32
    accumulation rounding (1e-5 of max |F|), for the MFMA and vector engines;
    J' differs from J by the e4m3 rounding (about 2^-5 relative)."""
    W = dense_normal_W(N, SEED_BASE + N)
    h, J, c = oracle.dense_hamiltonian(W.astype(np.float64))
    Jp, _ = fp8_ref.quantize(J.astype(np.float32).astype(np.float64))
    for engine in ("mfma", "vector"):
        an = make(rocm, W, m, field_precision="fp8", refresh_engine=engine)
        q = np.stack(an.get_q()).astype(np.float64)
        F = an.get_fields().astype(np.float64)
        ref = q @ Jp.T
        assert np.max(np.abs(F - ref)) <= 1e-5 * np.max(np.abs(ref)), engine
    exact = q @ J.T
    rel = np.max(np.abs(F - exact)) / np.max(np.abs(exact))
    assert 1e-4 < rel < 0.05


def test_form_energy_error(oracle, rocm):
    """get_form_energy_error = max_y |E_J'(q_y) - E_J(q_y)| in FP64."""
    N, m = 512, 8
    W = dense_normal_W(N, SEED_BASE + 9)
    h, J, c = oracle.dense_hamiltonian(W.astype(np.float64))
    J32 = J.astype(np.float32).astype(np.float64)
    Jp, _ = fp8_ref.quantize(J32)
    an = make(rocm, W, m, field_precision="fp8")
    for G in schedule()[:5]:
        an.anneal_one_step(G, BETA)
    q = np.stack(an.get_q()).astype(np.float64)
    h32 = h.astype(np.float32).astype(np.float64)

    def E(Jm):
        return -np.einsum("yi,ij,yj->y", q, Jm, q) - q @ h32

    ref = np.max(np.abs(E(Jp) - E(J32)))
    err = an.get_form_energy_error()
    assert ref > 0 and abs(err - ref) <= 1e-9 * max(1.0, ref)
    an32 = make(rocm, W, m, field_precision="fp32")
    assert an32.get_form_energy_error() == 0.0


@pytest.mark.parametrize("helper", ["on", "off"])
@pytest.mark.parametrize("N,m", [(128, 32), (1100, 6), (300, 7)])
def test_bit_exact_when_J_is_exact_in_e4m3(oracle, rocm, N, m, helper):
    """Integer W in [-8, 8] (J quarter-integral, |J| <= 2, exact in e4m3):
    the fp8 form's spins equal the serial CPU annealer's (MT19937 stream
    injected) after every 10th step, and the energies at the end.  With
    helper_wave=off the single-flip sweeps stream the e4m3 rows and their
    scales; with it on they run the helper-wave kernel on the float copy of
    J'."""
    W = dense_int_W(N, SEED_BASE + 5 * N + m)
    ora = oracle.DenseGraphAnnealer(np.float32, 1)
    ora.seed(1000 + N)
    ora.set_qubo(W)
    ora.set_m(m)
    ora.set_algorithm(3)
    ora.prepare()
    ora.randomize_spin()
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m, rng="mt19937", field_precision="fp8",
                                   helper_wave=helper)
    an.seed(1000 + N)
    an.prepare()
    an.randomize_spin()
    steps = schedule() if N <= 128 else schedule()[:40]
    for k, G in enumerate(steps):
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
        if k % 10 == 0 or k == len(steps) - 1:
            assert np.array_equal(np.stack(an.get_q()), ora.get_q()), f"step {k}"
    an.make_solution()
    assert np.array_equal(an.get_E(), ora.calculate_E())
    assert an.get_form_energy_error() == 0.0
