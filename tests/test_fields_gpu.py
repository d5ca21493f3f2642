"""Local fields of the dense annealer: the field GEMM F = Q.J on MFMA
(fp32, bf16 hi+lo and fp8) against the vector kernels and an exact/FP64 host
product, and the sweep's incremental row updates against a recompute."""
import numpy as np
import pytest

from problems import dense_int_W, dense_normal_W, schedule, SEED_BASE

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def annealer(rocm, W, m, **prefs):
    an = rocm.dense_graph_annealer(W, dtype=W.dtype.type, n_trotters=m, **prefs)
    an.seed(17)
    an.prepare()
    an.randomize_spin()
    return an


def host_fields(oracle, W, q):
    h, J, c = oracle.dense_hamiltonian(W.astype(np.float64))
    return q.astype(np.float64) @ J.T


@pytest.mark.parametrize("form", ["fp32", "bf16x2", "fp8"])
@pytest.mark.parametrize("N,m", [(300, 37), (128, 32), (1000, 16), (200, 520), (1100, 5)])
def test_field_gemm_exact_on_integer_problems(oracle, rocm, form, N, m):
    """Quarter-integer J (|J| <= 2, exact in every form): every engine gives
    the exact product (MFMA fp32, MFMA bf16 hi+lo, MFMA fp8, vector) --
    including ragged N and m, m > 256 (the MFMA kernels then take several
    passes over the slice rows) and an fp8 row stride (2048) above Np (1280)."""
    W = dense_int_W(N, SEED_BASE + N)
    res = {}
    for engine in ("mfma", "vector"):
        an = annealer(rocm, W, m, field_precision=form, refresh_engine=engine)
        res[engine] = an.get_fields().astype(np.float64)
        q = np.stack(an.get_q())
    exact = host_fields(oracle, W, q)
    assert np.array_equal(res["mfma"], exact)
    assert np.array_equal(res["vector"], exact)


def test_field_gemm_precision_forms(oracle, rocm):
    """W ~ N(0,1), N=2048, m=64: max |F - F_fp64| / max |F| is below 1e-5 for
    fp32 (J split exactly into three bf16 planes) and 2e-5 for bf16 hi+lo (J
    kept to 16 significant bits); fp32 accumulation dominates both."""
    N, m = 2048, 64
    W = dense_normal_W(N, SEED_BASE + 7)
    errs = {}
    for form in ("fp32", "bf16x2"):
        an = annealer(rocm, W, m, field_precision=form)
        F = an.get_fields().astype(np.float64)
        ref = host_fields(oracle, W, np.stack(an.get_q()))
        errs[form] = np.max(np.abs(F - ref)) / np.max(np.abs(ref))
    print("field GEMM relative errors:", errs)
    assert errs["fp32"] < 1e-5, errs
    assert errs["bf16x2"] < 2e-5, errs


@pytest.mark.parametrize("form", ["fp32", "bf16x2", "fp8"])
def test_incremental_fields_track_spins(oracle, rocm, form):
    """With refresh_interval=0 the sweep's row updates alone keep F = J.q
    exact for an integer problem over 40 annealing steps."""
    N, m = 512, 24
    W = dense_int_W(N, SEED_BASE + 31)
    an = annealer(rocm, W, m, field_precision=form, refresh_interval=0)
    for G in schedule()[:40]:
        an.anneal_one_step(G, 50.0)
    F = an.get_fields().astype(np.float64)
    assert np.array_equal(F, host_fields(oracle, W, np.stack(an.get_q())))
    assert an.get_counters()["flips"] > 0
