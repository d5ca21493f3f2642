"""The dense batch engine: instances of uneven N and m stepped together in
shared launches (grouped field GEMM, grouped Philox streams, one persistent
sweep, optionally one hipGraph per schedule) end bit-identical to the same
instances annealed alone, whatever the launch shape the autotuner picks."""
import numpy as np
import pytest

from problems import dense_int_W, dense_normal_W, schedule, SEED_BASE

pytestmark = pytest.mark.gpu

BETA = 50.0
SHAPES = [(64, 16), (300, 7), (1000, 32), (128, 2), (777, 5), (2048, 16), (40, 1)]


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def problems():
    out = []
    for i, (N, m) in enumerate(SHAPES):
        W = dense_int_W(N, SEED_BASE + 100 + i) if i % 2 == 0 else dense_normal_W(N, SEED_BASE + 100 + i)
        out.append((W, m, 5000 + 17 * i))
    return out


def alone(rocm, W, m, seed, steps, **prefs):
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m, **prefs)
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    for G in steps:
        an.anneal_one_step(G, BETA)
    return np.stack(an.get_q()), an.get_E()


def batch(rocm, probs, steps, graph=True, tune=False, **prefs):
    b = rocm.dense_batch(**prefs)
    for W, m, seed in probs:
        b.add_qubo(W, m, seed)
    b.prepare()
    if tune:
        b.autotune(steps[0], BETA, 1)
    b.anneal(steps, BETA, graph=graph)
    return b


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_instances_alone(rocm, graph):
    steps = schedule()[:25]
    probs = problems()
    b = batch(rocm, probs, steps, graph=graph)
    assert len(b) == len(probs)
    for i, (W, m, seed) in enumerate(probs):
        q, E = alone(rocm, W, m, seed, steps)
        assert np.array_equal(b.get_q(i), q), f"instance {i} (N={W.shape[0]}, m={m})"
        assert np.array_equal(b.get_E(i), E)


@pytest.mark.parametrize("form", ["bf16x2", "fp8"])
def test_batch_field_forms(rocm, form):
    steps = schedule()[:12]
    probs = problems()[:5]
    b = batch(rocm, probs, steps, field_precision=form)
    for i, (W, m, seed) in enumerate(probs):
        q, _ = alone(rocm, W, m, seed, steps, field_precision=form)
        assert np.array_equal(b.get_q(i), q), i


@pytest.mark.parametrize("form", ["fp32", "bf16x2"])
def test_batch_k_split_gemm_path(rocm, form):
    """An instance of 2500 sites (2560 padded, a k-range above 2048) takes the
    grouped field GEMM's k-split path (the block's four waves share the
    k-range and add their partial tiles in LDS) beside one that takes the
    column split; both end equal to the instances annealed alone."""
    steps = schedule()[:6]
    probs = [(dense_normal_W(2500, SEED_BASE + 140), 4, 7001), (dense_int_W(300, SEED_BASE + 141), 3, 7002)]
    b = batch(rocm, probs, steps, field_precision=form)
    for i, (W, m, seed) in enumerate(probs):
        q, E = alone(rocm, W, m, seed, steps, field_precision=form)
        assert np.array_equal(b.get_q(i), q), f"instance {i} (N={W.shape[0]}, m={m})"
        assert np.array_equal(b.get_E(i), E), i


def test_autotune_and_schedule_split_do_not_change_results(rocm):
    """Autotuning restores spins, fields and streams; a schedule run as two
    graphs equals one run as a single graph."""
    steps = schedule()[:20]
    probs = problems()
    ref = batch(rocm, probs, steps)
    tuned = batch(rocm, probs, steps, tune=True)
    b = rocm.dense_batch(order="given", waves_per_block=2)
    for W, m, seed in probs:
        b.add_qubo(W, m, seed)
    b.prepare()
    b.anneal(steps[:7], BETA)
    b.anneal(steps[7:], BETA)
    for i in range(len(probs)):
        assert np.array_equal(tuned.get_q(i), ref.get_q(i))
        assert np.array_equal(b.get_q(i), ref.get_q(i))


def test_batch_rejects_unbatchable_preferences(rocm):
    b = rocm.dense_batch()
    with pytest.raises(RuntimeError):
        b.set_preferences(rng="mt19937")
    with pytest.raises(RuntimeError):
        b.set_preferences(waves_per_block=3)
