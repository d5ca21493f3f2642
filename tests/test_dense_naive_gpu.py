"""The naive dense step on the rocm device (annealOneStepNaive,
sqaodc/cpu/CPUDenseGraphAnnealer.cpp:268-280: N m attempts, each drawing its
slice y = randInt(m) before tryFlip draws x and u) against the oracle's
restatement, bit for bit on integer problems: with rng = mt19937 the device
consumes the serial host generator's words in the reference's order (y, x,
u); with the default Philox stream the oracle generates the device's naive
layout itself."""
import numpy as np
import pytest

from problems import SEED_BASE, dense_int_W, schedule

pytestmark = pytest.mark.gpu

BETA = 50.0


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def pair(oracle, rocm, W, m, seed, rng, dtype=np.float32, opt=0, **prefs):
    import sqaod_amd as sq
    ora = oracle.DenseGraphAnnealer(dtype, 1)
    ora.seed(seed)
    ora.set_qubo(W.astype(dtype), opt)
    ora.set_m(m)
    ora.set_algorithm(oracle.ALGO_NAIVE)
    if rng == "philox":
        ora.set_stream(oracle.STREAM_PHILOX)
    ora.prepare()
    ora.randomize_spin()
    an = rocm.dense_graph_annealer(W.astype(dtype), sq.maximize if opt else sq.minimize, dtype=dtype,
                                   n_trotters=m, algorithm="naive", rng=rng, **prefs)
    assert an.get_preferences()["algorithm"] == "naive"
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    return ora, an


@pytest.mark.parametrize("rng", ["mt19937", "philox"])
@pytest.mark.parametrize("form", ["fp32", "bf16x2", "fp8"])
def test_naive_config1_bit_exact(oracle, rocm, rng, form):
    """N = 128, m = 32, integer W in [-8, 8], the 205-step schedule."""
    N, m = 128, 32
    W = dense_int_W(N, SEED_BASE)
    ora, an = pair(oracle, rocm, W, m, 1000, rng, field_precision=form)
    assert np.array_equal(np.stack(an.get_q()), ora.get_q())
    sch = schedule()
    for k, G in enumerate(sch):
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
        if k % 20 == 0 or k == len(sch) - 1:
            qa, qo = np.stack(an.get_q()), ora.get_q()
            assert np.array_equal(qa, qo), f"step {k}: {int(np.sum(qa != qo))} spins differ"
    an.make_solution()
    assert np.array_equal(an.get_E(), ora.calculate_E())


@pytest.mark.parametrize("N,m,opt", [(37, 2, 0), (300, 5, 1), (64, 3, 0)])
def test_naive_odd_shapes_and_maximize(oracle, rocm, N, m, opt):
    """m = 2 (both neighbours one slice), odd m, N not a multiple of 32,
    maximize; 40 steps with the MT19937 stream."""
    W = dense_int_W(N, SEED_BASE + 60 + N)
    ora, an = pair(oracle, rocm, W, m, 5 + m, "mt19937", opt=opt)
    for G in schedule()[:40]:
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
    assert np.array_equal(np.stack(an.get_q()), ora.get_q())
    assert np.array_equal(an.get_E(), ora.calculate_E())


def test_naive_after_plain_steps_and_set_q(oracle, rocm):
    """Switching algorithms keeps the trajectory: coloring steps with the
    plain sweep (which keeps no local fields), then naive steps, then set_q;
    the device recomputes the fields it needs."""
    N, m = 96, 4
    W = dense_int_W(N, SEED_BASE + 77)
    ora, an = pair(oracle, rocm, W, m, 21, "mt19937", sweep="plain")
    ora.set_algorithm(oracle.ALGO_COLORING)
    an.set_preferences(algorithm="coloring")
    for G in schedule()[:10]:
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
    ora.set_algorithm(oracle.ALGO_NAIVE)
    an.set_preferences(algorithm="naive")
    for G in schedule()[10:30]:
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
    assert np.array_equal(np.stack(an.get_q()), ora.get_q())
    q = np.where(np.arange(N) % 3 == 0, 1, -1).astype(np.int8)
    an.set_q(q)
    ora.set_q(np.tile(q, (m, 1)))
    for G in schedule()[30:50]:
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
    assert np.array_equal(np.stack(an.get_q()), ora.get_q())
