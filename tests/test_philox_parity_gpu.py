"""The default random stream (rng = philox) against the oracle.

The oracle generates the rocm device's Philox4x32-10 layouts itself
(oracle.STREAM_PHILOX; the layouts are pinned by the published Philox known
answers in tests/test_oracle_streams_cpu.py): attempt (y, k) of sweep s is the
block at counter (k, y, s_lo, s_hi) keyed by the annealer seed, site = word 0
% N and u = word 1 as Random.cpp:157-161 converts it, and randomize_spin
draws slice y's spins from the blocks at (x / 4, y, salt, ~0).  Fed that
stream, the oracle's restatement of sqaod.cpu (tryFlip,
CPUDenseGraphAnnealer.cpp:250-265, in the coloring order of :282-301; the
bipartite half-steps of CPUBipartiteGraphAnnealer.cpp:331-373) must give the
same spins, x and E as the device, bit for bit, on integer-valued problems
(every field and energy exact in float).  Covered here:

* the dense annealer in every field form (fp32, bf16x2, fp8, plain) and SA;
* the bipartite annealer (coloring and sa_coloring, both GEMM forms);
* all 512 instances of a batch with the batching benchmark's shape law
  (N log-uniform in [64, 2048], m in {16, 32, 64, 128}), integer W;
* a serving run with Poisson arrivals of mixed dense and bipartite jobs
  (N from 32 to 8192, m from 16 to 256, 50 to 500 steps each), every job
  checked against the oracle run alone.
"""
import numpy as np
import pytest

from problems import SEED_BASE, dense_int_W, schedule

pytestmark = pytest.mark.gpu

BETA = 50.0


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def int_bipartite(N0, N1, seed, lo=-8, hi=8):
    rng = np.random.default_rng(seed)
    b0 = rng.integers(lo, hi + 1, N0).astype(np.float32)
    b1 = rng.integers(lo, hi + 1, N1).astype(np.float32)
    W = rng.integers(lo, hi + 1, (N1, N0)).astype(np.float32)
    return b0, b1, W


def oracle_dense(oracle, W, m, seed, algo="coloring"):
    ora = oracle.DenseGraphAnnealer(np.float32, 1)
    ora.seed(seed)
    ora.set_qubo(W)
    ora.set_m(m)
    ora.set_algorithm({"coloring": oracle.ALGO_COLORING, "sa_naive": oracle.ALGO_SA_NAIVE}[algo])
    ora.set_stream(oracle.STREAM_PHILOX)
    ora.prepare()
    ora.randomize_spin()
    ora.set_local_fields(True)
    return ora


def oracle_bipartite(oracle, b0, b1, W, m, seed, algo="coloring"):
    ora = oracle.BipartiteGraphAnnealer(np.float32, 1)
    ora.seed(seed)
    ora.set_qubo(b0, b1, W)
    ora.set_m(m)
    ora.set_algorithm(oracle.ALGO_SA_COLORING if algo == "sa_coloring" else oracle.ALGO_COLORING)
    ora.set_stream(oracle.STREAM_PHILOX)
    ora.prepare()
    ora.randomize_spin()
    return ora


def pair_arrays(qpairs):
    return np.stack([p[0] for p in qpairs]), np.stack([p[1] for p in qpairs])


@pytest.mark.parametrize("prefs", [dict(field_precision="fp32"), dict(field_precision="bf16x2"),
                                   dict(field_precision="fp8"), dict(sweep="plain")],
                         ids=["fp32", "bf16x2", "fp8", "plain"])
def test_dense_default_stream_config1(oracle, rocm, prefs):
    """N = 128, m = 32, W integer in [-8, 8], the full 205-step schedule: the
    default stream's trajectory equals the oracle's after every 10th step."""
    N, m, seed = 128, 32, 1000
    W = dense_int_W(N, SEED_BASE)
    ora = oracle_dense(oracle, W, m, seed)
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m, **prefs)
    assert an.get_preferences()["rng"] == "philox"
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    assert np.array_equal(np.stack(an.get_q()), ora.get_q()), "randomize_spin differs"
    sch = schedule()
    for k, G in enumerate(sch):
        an.anneal_one_step(G, BETA)
        ora.anneal_one_step(G, BETA)
        if k % 10 == 0 or k == len(sch) - 1:
            qa, qo = np.stack(an.get_q()), ora.get_q()
            assert np.array_equal(qa, qo), f"step {k}: {int(np.sum(qa != qo))} spins differ"
    an.make_solution()
    assert np.array_equal(an.get_E(), ora.calculate_E())
    assert np.array_equal(np.stack(an.get_x()), (ora.get_q() + 1) // 2)


@pytest.mark.parametrize("N,m,form", [(1500, 64, "bf16x2"), (1500, 64, "fp8"), (777, 7, "fp32"),
                                      (2048, 256, "bf16x2")])
def test_dense_default_stream_larger(oracle, rocm, N, m, form):
    """Larger and odd shapes (N not a multiple of 256, odd m, m = 256), 40
    steps from the start of the schedule (the flip-heavy part)."""
    seed = 17 + m
    W = dense_int_W(N, SEED_BASE + N + m)
    ora = oracle_dense(oracle, W, m, seed)
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m, field_precision=form)
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    for k, G in enumerate(schedule()[:40]):
        an.anneal_one_step(G, BETA)
        ora.anneal_one_step(G, BETA)
        if k % 13 == 0:
            assert np.array_equal(np.stack(an.get_q()), ora.get_q()), f"step {k}"
    assert np.array_equal(np.stack(an.get_q()), ora.get_q())
    assert np.array_equal(an.get_E(), ora.calculate_E())


def test_dense_default_stream_sa(oracle, rocm):
    """sa_naive (CPUDenseGraphAnnealer.cpp:341-371), kT = G, slices in turn."""
    N, m, seed = 200, 3, 5
    W = dense_int_W(N, SEED_BASE + 31)
    ora = oracle_dense(oracle, W, m, seed, algo="sa_naive")
    an = rocm.dense_graph_annealer(W, dtype=np.float32, n_trotters=m, algorithm="sa_naive")
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    for kT in np.geomspace(30.0, 0.05, 50):
        an.anneal_one_step(kT, 0.0)
        ora.anneal_one_step(kT, 0.0)
    assert np.array_equal(np.stack(an.get_q()), ora.get_q())


@pytest.mark.parametrize("N0,N1,m,algo,form", [(300, 200, 16, "coloring", "fp32"),
                                               (300, 200, 16, "coloring", "bf16x2"),
                                               (129, 64, 5, "coloring", "fp32"),
                                               (100, 160, 4, "sa_coloring", "fp32")])
def test_bipartite_default_stream(oracle, rocm, N0, N1, m, algo, form):
    b0, b1, W = int_bipartite(N0, N1, SEED_BASE + 900 + N0 + m)
    seed = 23 + m
    an = rocm.bipartite_graph_annealer(b0, b1, W, dtype=np.float32, n_trotters=m, algorithm=algo,
                                       field_precision=form)
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    ora = oracle_bipartite(oracle, b0, b1, W, m, seed, algo)
    q0, q1 = pair_arrays(an.get_q())
    r0, r1 = ora.get_q()
    assert np.array_equal(q0, r0) and np.array_equal(q1, r1), "randomize_spin differs"
    steps = schedule()[:80] if algo == "coloring" else [5.0 * 0.93 ** k for k in range(60)]
    for k, G in enumerate(steps):
        an.anneal_one_step(G, BETA)
        ora.anneal_one_step(G, BETA)
        if k % 9 == 0 or k == len(steps) - 1:
            q0, q1 = pair_arrays(an.get_q())
            r0, r1 = ora.get_q()
            assert np.array_equal(q0, r0) and np.array_equal(q1, r1), f"step {k}"
    assert np.array_equal(an.get_E(), ora.calculate_E())


def batch_shapes(n, seed):
    """The batching benchmark's shape law: N = round(64 * 32^u), u ~ U[0, 1),
    m uniform in {16, 32, 64, 128}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = int(round(64 * 32 ** rng.random()))
        m = int(rng.choice([16, 32, 64, 128]))
        out.append((N, m))
    return out


def test_batch_512_instances_against_oracle(oracle, rocm):
    """All 512 instances of a batch (one persistent sweep per step, the
    schedule's first steps captured as one hipGraph) against the oracle run
    alone, instance by instance: spins after the steps and the energies."""
    shapes = batch_shapes(512, 4242)
    steps = schedule()[:3]
    b = rocm.dense_batch(field_precision="bf16x2")
    Ws = []
    for i, (N, m) in enumerate(shapes):
        W = dense_int_W(N, SEED_BASE + 7000 + i)
        Ws.append(W)
        assert b.add_qubo(W, m, 1000 + i) == i
    b.prepare()
    b.anneal(steps, BETA, graph=True)
    bad = []
    for i, (N, m) in enumerate(shapes):
        ora = oracle_dense(oracle, Ws[i], m, 1000 + i)
        for G in steps:
            ora.anneal_one_step(G, BETA)
        if not np.array_equal(b.get_q(i), ora.get_q()) or not np.array_equal(b.get_E(i), ora.calculate_E()):
            bad.append((i, N, m))
    assert not bad, f"{len(bad)} of 512 instances differ from the oracle, first {bad[:5]}"


def serve_jobs():
    """A scaled serving mix: 70% dense / 30% bipartite (N0 = N1 = N / 2), N
    from 32 to 8192, m from 16 to 256, 50 to 500 steps with a per-job
    geometric G schedule and beta; integer weights in [-8, 8]."""
    spec = [("dense", 32, 16, 500), ("dense", 100, 256, 120), ("dense", 700, 64, 300),
            ("dense", 2048, 200, 60), ("dense", 8192, 16, 50), ("dense", 3000, 128, 50),
            ("dense", 250, 40, 200), ("bipartite", 64, 32, 400), ("bipartite", 1024, 64, 80),
            ("bipartite", 4096, 16, 50)]
    rng = np.random.default_rng(SEED_BASE + 555)
    jobs = []
    for i, (kind, N, m, n) in enumerate(spec):
        G0, beta = float(rng.uniform(2.0, 5.0)), float(rng.uniform(20.0, 60.0))
        G = G0 * (0.01 / G0) ** (np.arange(n) / (n - 1))
        if kind == "dense":
            data = (dense_int_W(N, SEED_BASE + 8000 + i),)
        else:
            data = int_bipartite(N // 2, N // 2, SEED_BASE + 8000 + i)
        jobs.append((kind, data, m, 5000 + i, G, beta))
    return jobs


def test_serving_poisson_mix_against_oracle(oracle, rocm):
    jobs = serve_jobs()
    srv = rocm.server()
    rng = np.random.default_rng(99)
    arrivals = np.cumsum(rng.exponential(4.0, len(jobs))).astype(int)  # in ticks
    ids, tick, nxt = {}, 0, 0
    while True:
        while nxt < len(jobs) and arrivals[nxt] <= tick:
            kind, data, m, seed, G, beta = jobs[nxt]
            bvec = np.full(len(G), beta)
            if kind == "dense":
                ids[srv.submit_dense(data[0], m, seed, G, bvec)] = nxt
            else:
                ids[srv.submit_bipartite(*data, m, seed, G, bvec)] = nxt
            nxt += 1
        left = srv.tick()
        tick += 1
        if left == 0 and nxt >= len(jobs):
            break
    results = {}
    for d in srv.take_done():
        results[ids[d]] = srv.result(d)
        srv.release(d)
    assert sorted(results) == list(range(len(jobs)))
    for i, (kind, data, m, seed, G, beta) in enumerate(jobs):
        r = results[i]
        if kind == "dense":
            ora = oracle_dense(oracle, data[0], m, seed)
        else:
            ora = oracle_bipartite(oracle, *data, m, seed)
            ora.set_local_fields(True)
        for g in G:
            ora.anneal_one_step(g, beta)
        E = ora.calculate_E()
        assert np.array_equal(r["E"], E), f"job {i} ({kind}, N={data[0].shape[0]}, m={m}): E differs"
        if kind == "dense":
            assert np.array_equal(r["x"], (ora.get_q() + 1) // 2), f"job {i}: x differs"
        else:
            q0, q1 = ora.get_q()
            assert np.array_equal(r["x0"], (q0 + 1) // 2) and np.array_equal(r["x1"], (q1 + 1) // 2), f"job {i}"
