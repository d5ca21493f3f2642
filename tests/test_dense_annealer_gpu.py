"""GPU parity of the rocm dense annealer against the CPU oracle.

Bit-exact cases use integer-valued W (J quarter-integral, every field and
energy exact in float) and the device MT19937 stream (rng=mt19937), which
replays the serial CPU annealer's random stream; the oracle is the C
restatement of CPUDenseGraphAnnealer (oracle/).  Float cases compare energies
with an FP64 recompute within stated tolerances.
"""
import numpy as np
import pytest

from problems import dense_int_W, dense_normal_W, schedule, SEED_BASE, dense_graph_8x8

pytestmark = pytest.mark.gpu

BETA = 50.0


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def make_pair(oracle, rocm, W, m, seed, algo="coloring", dtype=np.float32, **prefs):
    N = W.shape[0]
    ora = oracle.DenseGraphAnnealer(dtype, 1)
    ora.seed(seed)
    ora.set_qubo(W.astype(dtype))
    ora.set_m(m)
    ora.set_algorithm({"coloring": 3, "sa_naive": 6}[algo])
    ora.prepare()
    ora.randomize_spin()
    an = rocm.dense_graph_annealer(W.astype(dtype), dtype=dtype, n_trotters=m, algorithm=algo,
                                   rng="mt19937", **prefs)
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    return ora, an


def q_of(an):
    return np.stack(an.get_q())


@pytest.mark.parametrize("variant", [dict(), dict(field_precision="bf16x2"), dict(field_precision="fp8"),
                                     dict(sweep="plain"), dict(slices_per_block=1),
                                     dict(sweep_launch="cooperative")])
def test_config1_bit_exact_trajectory(oracle, rocm, variant):
    """N=128, m=32, integer W in [-8, 8], G 5.0 -> 0.01 (x0.97), beta=50: the
    spins of every slice and the energies equal the serial CPU annealer's
    after every step of the schedule."""
    N, m = 128, 32
    W = dense_int_W(N, SEED_BASE)
    ora, an = make_pair(oracle, rocm, W, m, seed=1000, **variant)
    assert np.array_equal(q_of(an), ora.get_q()), "randomize_spin differs"
    for step, G in enumerate(schedule()):
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
        if step % 10 == 0 or step == len(schedule()) - 1:
            qa, qo = q_of(an), ora.get_q()
            assert np.array_equal(qa, qo), f"step {step}: {np.sum(qa != qo)} spins differ"
    an.make_solution()
    E = an.get_E()
    assert np.array_equal(E, ora.calculate_E())
    x = np.stack(an.get_x())
    assert np.array_equal(x, (ora.get_q() + 1) // 2)


@pytest.mark.parametrize("N,m", [(100, 5), (64, 2), (37, 3), (200, 7), (256, 16), (1, 1), (1, 3), (2, 1),
                                 (33, 1)])
def test_odd_shapes_bit_exact(oracle, rocm, N, m):
    """odd m (ring neighbours 0 and m-1 share a colour), m = 2 (both
    neighbours one slice), m = 1 (the slice is its own neighbour), N = 1 and N
    not a multiple of 32: bit-exact over 30 steps."""
    W = dense_int_W(N, SEED_BASE + N + m)
    ora, an = make_pair(oracle, rocm, W, m, seed=7 + m)
    for G in schedule()[:30]:
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
    assert np.array_equal(q_of(an), ora.get_q())


def test_sa_naive_bit_exact(oracle, rocm):
    """m = 1 selects SA (CPUDenseGraphAnnealer.cpp:174-175); kT = G."""
    N, m = 96, 1
    W = dense_int_W(N, SEED_BASE + 5)
    ora, an = make_pair(oracle, rocm, W, m, seed=3, algo="sa_naive")
    assert an.get_preferences()["algorithm"] == "sa_naive"
    for kT in np.geomspace(20.0, 0.05, 60):
        ora.anneal_one_step(kT, 0.0)
        an.anneal_one_step(kT, 0.0)
    assert np.array_equal(q_of(an), ora.get_q())


def test_maximize_bit_exact(oracle, rocm):
    N, m = 64, 8
    W = dense_int_W(N, SEED_BASE + 9)
    ora = oracle.DenseGraphAnnealer(np.float32, 1)
    ora.seed(11)
    ora.set_qubo(W, oracle.MAXIMIZE)
    ora.set_m(m)
    ora.prepare()
    ora.randomize_spin()
    import sqaod_amd as sq
    an = rocm.dense_graph_annealer(W, sq.maximize, dtype=np.float32, n_trotters=m, rng="mt19937")
    an.seed(11)
    an.prepare()
    an.randomize_spin()
    for G in schedule()[:40]:
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
    assert np.array_equal(q_of(an), ora.get_q())
    assert np.array_equal(an.get_E(), ora.calculate_E())


def test_float_energy_fp64(oracle, rocm):
    """W ~ N(0,1): get_E is the FP64 recompute rounded to float; it agrees
    with the oracle's FP64 energy of the same spins within 1e-5 relative."""
    N, m = 512, 16
    W = dense_normal_W(N, SEED_BASE + 2)
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m)
    an.seed(5)
    an.prepare()
    an.randomize_spin()
    for G in schedule()[:50]:
        an.anneal_one_step(G, BETA)
    E = an.get_E()
    q = np.stack(an.get_q())
    h, J, c = oracle.dense_hamiltonian(W.astype(np.float64))
    Eref = oracle.dense_energy_q(h, J, c, q)
    assert np.allclose(E, Eref, rtol=1e-5, atol=1e-3)


def test_local_vs_plain_float(rocm):
    """The local-field sweep and the plain per-attempt-dot sweep make the same
    decisions on float data with the same stream (fields agree to rounding)."""
    N, m = 256, 8
    W = dense_normal_W(N, SEED_BASE + 3)
    res = []
    for sweep in ("local", "plain"):
        an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m, sweep=sweep)
        an.seed(21)
        an.prepare()
        an.randomize_spin()
        for G in schedule()[:20]:
            an.anneal_one_step(G, BETA)
        res.append(np.stack(an.get_q()))
    frac = np.mean(res[0] != res[1])
    assert frac < 0.05


def test_known_answers(rocm):
    """sqaodpy/tests/test_dense_graph_annealer.py:85-96 and :154-178."""
    N = 8
    an = rocm.dense_graph_annealer(np.ones((N, N), np.float32), dtype=np.float32, n_trotters=1)
    an.prepare()
    an.set_q(np.ones(N, np.int8))
    E = an.get_E()
    assert len(E) == 1 and E[0] == N * N
    x = an.get_x()
    assert len(x) == 1 and np.all(x[0] == 1)
    an.set_q(-np.ones(N, np.int8))
    assert an.get_E()[0] == 0.


def test_trivial_problem_reaches_zero(rocm):
    """test_dense_graph_annealer.py:180-260: W = ones(10); minimize reaches E=0."""
    import sqaod_amd as sq
    for opt, sign in ((sq.minimize, 1), (sq.maximize, -1)):
        W = sign * np.ones((10, 10), np.float64)
        an = rocm.dense_graph_annealer(W, opt, dtype=np.float64, n_trotters=4)
        an.seed(0)
        an.prepare()
        an.randomize_spin()
        G = 5.0
        while G > 0.02:
            an.anneal_one_step(G, 1. / 0.03)
            G *= 0.95
        an.make_solution()
        E = an.get_E()
        assert np.min(E) == 0 if opt is sq.minimize else np.max(E) == 0


def test_8x8_example(rocm, oracle):
    W = dense_graph_8x8(np.float32)
    Emin, _ = oracle.dense_bf_search(W.astype(np.float64))
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=4)
    an.seed(1)
    an.prepare()
    an.randomize_spin()
    for G in schedule():
        an.anneal_one_step(G, BETA)
    assert np.min(an.get_E()) == Emin


def test_large_instance_runs(rocm, oracle):
    """N=2048, m=64, W ~ N(0,1): the sweep completes (no neighbour-wait
    abort), the energy decreases, and E agrees with an FP64 recompute."""
    N, m = 2048, 64
    W = dense_normal_W(N, SEED_BASE + 4)
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m)
    an.seed(3)
    an.prepare()
    an.randomize_spin()
    E0 = np.mean(an.get_E())
    for G in schedule()[:40]:
        an.anneal_one_step(G, BETA)
    E1 = an.get_E()
    assert np.mean(E1) < E0
    q = np.stack(an.get_q())
    h, J, c = oracle.dense_hamiltonian(W.astype(np.float64))
    Eref = oracle.dense_energy_q(h, J, c, q)
    assert np.allclose(E1, Eref, rtol=1e-5, atol=1e-2)
    cnt = an.get_counters()
    assert cnt["flips"] > 0 and cnt["batches"] > 0


@pytest.mark.parametrize("form", ["fp32", "bf16x2", "fp8"])
def test_largest_lds_resident_n_bit_exact(oracle, rocm, form):
    """N = 11520, the largest padded N whose float local-field sweep fits one
    workgroup's LDS (fields, staged row, pending table, spin bits, neighbour
    mirrors and windows: 12.4 bytes per site plus fixed parts, within 160 KiB),
    m = 4: spins equal the serial CPU annealer's after three steps.  The row
    DMA then spans 45 1-KiB ops (46 counted ops after the prefetch group) and
    the pending table is 4096 entries for 11520 sites."""
    N, m = 11520, 4
    W = dense_int_W(N, SEED_BASE + 11)
    ora, an = make_pair(oracle, rocm, W, m, seed=321, field_precision=form)
    for G in schedule()[:3]:
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
        qa, qo = q_of(an), ora.get_q()
        assert np.array_equal(qa, qo), f"{np.sum(qa != qo)} spins differ"


def test_spin_limit_gives_up_cleanly(rocm):
    """With a spin limit of one blocked iteration (forced from the host) the
    sweep's bounded wait gives up instead of waiting: the step raises
    RuntimeError, every kernel has exited, and the annealer works again once
    its spins are set anew."""
    W = dense_int_W(128, SEED_BASE + 41)
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=2, spin_limit=1)
    an.seed(3)
    an.prepare()
    an.randomize_spin()
    with pytest.raises(RuntimeError, match="spin limit"):
        for G in schedule()[:40]:
            an.anneal_one_step(G, BETA)
        an.get_q()  # a flag written by the last step is reported here at the latest
    with pytest.raises(RuntimeError):
        an.set_preferences(spin_limit=0)
    an.set_preferences(spin_limit=1 << 20)
    an.randomize_spin()
    for G in schedule()[:10]:
        an.anneal_one_step(G, BETA)
    an.make_solution()
    x = np.stack(an.get_x()).astype(np.float64)
    assert np.array_equal(an.get_E(), np.einsum("yi,yi->y", x @ W.astype(np.float64), x).astype(np.float32))


@pytest.mark.parametrize("T,N,m", [(8, 128, 32), (3, 128, 32), (16, 100, 7), (4, 64, 5), (2, 50, 2)])
def test_openmp_worker_streams_bit_exact(oracle, rocm, T, N, m):
    """sqaod.cpu's default on a multi-core host is the OpenMP coloring step
    (CPUDenseGraphAnnealer.cpp:303-339) with one MT19937 per worker seeded
    s + 17 w (:29-33) and the slices of each colour in static chunks; with
    n_workers = T the device replays that trajectory (the oracle's
    restatement with T workers): spins after every 10th step of the
    schedule, and the energies, bit for bit.  Odd m puts slice m-1 on worker
    0 between the colours; T above m/2 leaves workers without slices."""
    W = dense_int_W(N, SEED_BASE + 90 + T + m)
    ora = oracle.DenseGraphAnnealer(np.float32, T)
    ora.seed(1000)
    ora.set_qubo(W)
    ora.set_m(m)
    ora.prepare()
    ora.randomize_spin()
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m, rng="mt19937", n_workers=T)
    assert an.get_preferences()["n_workers"] == T
    an.seed(1000)
    an.prepare()
    an.randomize_spin()
    assert np.array_equal(q_of(an), ora.get_q())
    for k, G in enumerate(schedule()[:100]):
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
        if k % 10 == 0:
            assert np.array_equal(q_of(an), ora.get_q()), f"step {k}"
    assert np.array_equal(q_of(an), ora.get_q())
    assert np.array_equal(an.get_E(), ora.calculate_E())


def test_generators_continue_across_prepare(oracle, rocm):
    """prepare() called again does not restart the random streams (the CPU
    annealer seeds only in seed(), CPUDenseGraphAnnealer.cpp:29-33, 164-167):
    the reference tests' calling sequence prepare / randomize_spin / anneal,
    twice (test_dense_graph_annealer.py:356-361), stays bit-exact."""
    N, m = 60, 6
    W = dense_int_W(N, SEED_BASE + 97)
    ora, an = make_pair(oracle, rocm, W, m, seed=4)
    for rep in range(2):
        if rep:
            ora.prepare()
            ora.randomize_spin()
            an.prepare()
            an.randomize_spin()
        for G in schedule()[:15]:
            ora.anneal_one_step(G, BETA)
            an.anneal_one_step(G, BETA)
        assert np.array_equal(q_of(an), ora.get_q()), rep


@pytest.mark.parametrize("N,m,form,algo", [(128, 32, "fp32", "coloring"), (128, 32, "bf16x2", "coloring"),
                                           (128, 32, "fp8", "coloring"),
                                           (300, 7, "fp32", "coloring"), (64, 2, "fp32", "coloring"),
                                           (1000, 16, "bf16x2", "coloring"), (96, 1, "fp32", "sa_naive"),
                                           (2048, 64, "fp32", "coloring"), (5000, 4, "fp32", "coloring"),
                                           (6000, 3, "bf16x2", "coloring")])
def test_multi_flip_batches_bit_exact(oracle, rocm, N, m, form, algo):
    """multi_flip: a batch resolves all its accepted flips in order (each
    flip corrects the later lanes' fields with the gathered submatrix and the
    rows are applied in flip order), so the trajectory is still the serial
    CPU annealer's, bit for bit, in the flip-heavy start of the schedule and
    to its end.  N = 5000 and 6000 leave LDS for three and two row stage
    buffers (four at N <= 4096), so batches with more flips than buffers
    apply their earlier rows at once and defer fewer."""
    W = dense_int_W(N, SEED_BASE + 3 * N + m)
    # threshold 2: every batch after one with two accepting lanes is a multi-flip batch
    ora, an = make_pair(oracle, rocm, W, m, seed=31 + m, algo=algo, field_precision=form, multi_flip="2")
    assert an.get_preferences()["multi_flip"] == "2"
    sch = schedule() if N <= 300 else schedule()[:40]
    if algo == "sa_naive":
        sch = list(np.geomspace(20.0, 0.05, 60))
    for k, G in enumerate(sch):
        beta = 0.0 if algo == "sa_naive" else BETA
        ora.anneal_one_step(G, beta)
        an.anneal_one_step(G, beta)
        if k % 10 == 0 or k == len(sch) - 1:
            qa, qo = q_of(an), ora.get_q()
            assert np.array_equal(qa, qo), f"step {k}: {int(np.sum(qa != qo))} spins differ"
    assert np.array_equal(an.get_E(), ora.calculate_E())
    c = an.get_counters()
    assert c["flips"] > 0


@pytest.mark.parametrize("N,m", [(5000, 4), (6000, 3), (2048, 16)])
def test_multi_flip_engages_and_matches_single_flip(rocm, N, m):
    """The multi-flip sweep takes over where LDS holds at least two row stage
    buffers (four at N = 2048, three at 5000, two at 6000): in the flip-heavy
    first steps it commits the same spins as the single-flip sweep in fewer
    batches."""
    W = dense_int_W(N, SEED_BASE + 7 * N)
    runs = {}
    for mode in ("off", "2"):
        an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m, multi_flip=mode)
        an.seed(5)
        an.prepare()
        an.randomize_spin()
        for G in schedule()[:3]:
            an.anneal_one_step(G, BETA)
        runs[mode] = (q_of(an), an.get_counters())
    assert np.array_equal(runs["off"][0], runs["2"][0])
    assert runs["2"][1]["flips"] == runs["off"][1]["flips"]
    assert runs["2"][1]["batches"] < 0.8 * runs["off"][1]["batches"]


@pytest.mark.parametrize("N,m,form,algo", [(128, 32, "fp32", "coloring"), (300, 7, "fp32", "coloring"),
                                           (1000, 16, "bf16x2", "coloring"), (96, 1, "fp32", "sa_naive"),
                                           (4096, 8, "fp32", "coloring"), (6000, 3, "bf16x2", "coloring")])
def test_helper_wave_bit_exact(oracle, rocm, N, m, form, algo):
    """helper_wave=on (the default): a second wave per slice adds the last
    flip's J row to the LDS fields while the sweeping wave runs the next
    batch's neighbour phase, the next row streaming into the other stage
    buffer; the sweeping wave waits for it before reading a field.  With
    multi_flip=off every sweep runs that kernel, and the trajectory is still
    the serial CPU annealer's, bit for bit."""
    W = dense_int_W(N, SEED_BASE + 5 * N + m)
    ora, an = make_pair(oracle, rocm, W, m, seed=41 + m, algo=algo, field_precision=form, multi_flip="off",
                        helper_wave="on")
    assert an.get_preferences()["helper_wave"] == "on"
    sch = schedule() if N <= 300 else schedule()[:30]
    if algo == "sa_naive":
        sch = list(np.geomspace(20.0, 0.05, 60))
    for k, G in enumerate(sch):
        beta = 0.0 if algo == "sa_naive" else BETA
        ora.anneal_one_step(G, beta)
        an.anneal_one_step(G, beta)
        if k % 10 == 0 or k == len(sch) - 1:
            qa, qo = q_of(an), ora.get_q()
            assert np.array_equal(qa, qo), f"step {k}: {int(np.sum(qa != qo))} spins differ"
    assert np.array_equal(an.get_E(), ora.calculate_E())


def test_helper_wave_on_off_same_sweep(rocm):
    """The helper wave changes who adds a row to the fields, not the sweep:
    the same spins, flips and energies with it on and off (float normal W, so
    the fields carry rounding that any change of operation order would
    show).  Batch counts are not compared: where a batch ends depends on how
    far the neighbour slices have got, which is timing."""
    W = dense_normal_W(2048, SEED_BASE + 99)
    runs = {}
    for mode in ("off", "on"):
        an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=32, helper_wave=mode)
        an.seed(9)
        an.prepare()
        an.randomize_spin()
        for G in schedule()[:12]:
            an.anneal_one_step(G, BETA)
        runs[mode] = (q_of(an), an.get_counters(), an.get_E())
    assert np.array_equal(runs["off"][0], runs["on"][0])
    assert runs["off"][1]["flips"] == runs["on"][1]["flips"]
    assert np.array_equal(runs["off"][2], runs["on"][2])


@pytest.mark.parametrize("spb", [2, 4])
def test_helper_wave_with_several_slices_per_block(oracle, rocm, spb):
    """The helper wave needs one slice per workgroup: with slices_per_block
    2 or 4 and helper_wave=on the sweep runs the one-wave kernel, still the
    serial CPU annealer's trajectory."""
    N, m = 300, 7
    W = dense_int_W(N, SEED_BASE + 17 * spb)
    ora, an = make_pair(oracle, rocm, W, m, seed=61 + spb, slices_per_block=spb, helper_wave="on")
    assert int(an.get_preferences()["slices_per_block"]) == spb
    for k, G in enumerate(schedule()[:30]):
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
    assert np.array_equal(q_of(an), ora.get_q())
    assert np.array_equal(an.get_E(), ora.calculate_E())


def test_helper_wave_near_the_lds_limit(oracle, rocm):
    """N = 8192: fields, two row stage buffers, h and the pending table fill
    most of a workgroup's 160 KiB of LDS, and the helper-wave kernel still
    runs the serial CPU annealer's trajectory."""
    N, m = 8192, 2
    W = dense_int_W(N, SEED_BASE + 8192)
    ora, an = make_pair(oracle, rocm, W, m, seed=73, multi_flip="off", helper_wave="on")
    for G in schedule()[:8]:
        ora.anneal_one_step(G, BETA)
        an.anneal_one_step(G, BETA)
    assert np.array_equal(q_of(an), ora.get_q())
    assert np.array_equal(an.get_E(), ora.calculate_E())
