"""Serving with continuous batching: jobs of mixed kind and size arrive while
others run, are admitted in arrival order under a memory budget, and each
finished job's E and x equal the same job run alone (dense and bipartite
annealers with the same seed and schedule)."""
import numpy as np
import pytest

from problems import dense_int_W, dense_normal_W, SEED_BASE

pytestmark = pytest.mark.gpu

BETA = 40.0


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


def sched(G0, n):
    return G0 * 0.95 ** np.arange(n)


def jobs():
    rng = np.random.default_rng(SEED_BASE + 77)
    out = []
    for i, (N, m, n) in enumerate([(64, 16, 30), (300, 7, 41), (1000, 32, 25), (2048, 16, 12), (5000, 8, 6),
                                   (129, 1, 20)]):
        W = dense_int_W(N, SEED_BASE + 300 + i) if i % 2 == 0 else dense_normal_W(N, SEED_BASE + 300 + i)
        out.append(("dense", (W,), m, 900 + i, sched(3.0 + i, n)))
    for i, (N0, N1, m, n) in enumerate([(100, 80, 6, 33), (512, 512, 16, 15), (37, 300, 5, 22)]):
        b0 = rng.standard_normal(N0).astype(np.float32)
        b1 = rng.standard_normal(N1).astype(np.float32)
        W = rng.standard_normal((N1, N0)).astype(np.float32)
        out.append(("bipartite", (b0, b1, W), m, 700 + i, sched(2.0 + i, n)))
    return out


def alone(rocm, kind, data, m, seed, G):
    if kind == "dense":
        an = rocm.dense_graph_annealer(data[0], dtype=np.float32, n_trotters=m, field_precision="bf16x2")
    else:
        an = rocm.bipartite_graph_annealer(*data, dtype=np.float32, n_trotters=m, field_precision="bf16x2")
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    for g in G:
        an.anneal_one_step(g, BETA)
    E = an.get_E()
    x = an.get_x()
    if kind == "dense":
        return E, (np.stack(x),)
    return E, (np.stack([p[0] for p in x]), np.stack([p[1] for p in x]))


def submit(srv, job):
    kind, data, m, seed, G = job
    if kind == "dense":
        return srv.submit_dense(data[0], m, seed, G, BETA)
    return srv.submit_bipartite(*data, m, seed, G, BETA)


def test_results_equal_jobs_run_alone(rocm):
    """Jobs arrive in three waves while earlier ones are running."""
    js = jobs()
    srv = rocm.server(dense_field_precision="bf16x2", bipartite_field_precision="bf16x2", bipartite_streams=2)
    ids = {}
    waves = [js[0:3] + js[6:7], js[3:5] + js[7:8], js[5:6] + js[8:9]]
    results = {}
    for w in waves:
        for j in w:
            ids[submit(srv, j)] = j
        for _ in range(4):
            srv.tick()
            for d in srv.take_done():
                results[d] = srv.result(d)
    while srv.tick() > 0:
        for d in srv.take_done():
            results[d] = srv.result(d)
    for d in srv.take_done():
        results[d] = srv.result(d)
    assert sorted(results) == sorted(ids)
    for jid, (kind, data, m, seed, G) in ids.items():
        r = results[jid]
        E, xs = alone(rocm, kind, data, m, seed, G)
        assert np.array_equal(r["E"], E), (jid, kind)
        if kind == "dense":
            assert np.array_equal(r["x"], xs[0]), jid
        else:
            assert np.array_equal(r["x0"], xs[0]) and np.array_equal(r["x1"], xs[1]), jid
        assert r["submit_s"] <= r["admit_s"] <= r["finish_s"]


def test_memory_budget_admission_in_arrival_order(rocm):
    """A budget that holds only some of the jobs at a time (dense N = 64, 300,
    1000 take about 28 MiB by the admission estimate, N = 2048 about 78 MiB,
    the budget is 90 MiB): the others wait, the estimate in use never exceeds
    the budget, admission follows arrival order (the last small job waits
    behind the large one), and every job completes."""
    js = jobs()[:4] + jobs()[:1]
    srv = rocm.server(memory_budget_mb=90, max_active=3)
    ids = [submit(srv, j) for j in js]
    peak_active = 0
    while True:
        rem = srv.tick()
        st = srv.status()
        assert st["bytes_in_use"] <= 90 << 20
        peak_active = max(peak_active, st["active"])
        if rem == 0:
            break
    done = srv.take_done()
    assert sorted(done) == sorted(ids)
    assert peak_active <= 3
    admits = [srv.result(i)["admit_s"] for i in ids]
    assert admits == sorted(admits)
    assert srv.status()["peak_bytes"] <= 90 << 20
    r = [srv.result(i) for i in ids]
    # N = 2048 fits only once N = 1000 (the first job to finish) has left
    assert r[3]["admit_s"] >= r[2]["finish_s"]
    assert r[4]["admit_s"] >= r[3]["admit_s"]


def test_rejects_job_over_budget_and_bad_input(rocm):
    srv = rocm.server(memory_budget_mb=8)
    with pytest.raises(RuntimeError):
        srv.submit_dense(dense_int_W(2048, 1), 64, 1, sched(1.0, 3), BETA)
    with pytest.raises(RuntimeError):
        srv.submit_dense(np.arange(16, dtype=np.float32).reshape(4, 4), 4, 1, sched(1.0, 3), BETA)
    with pytest.raises(RuntimeError):
        srv.set_preferences(dense_field_precision="fp16")


def test_rejects_dense_job_the_grouped_sweep_cannot_hold(rocm):
    """A dense job whose slices cannot all be resident (m above the waves the
    sweep may use) or whose fields do not fit a wave's LDS region (N above
    the LDS-resident limit) is rejected at submission instead of failing
    every tick."""
    srv = rocm.server(block_fraction=0.001)  # leaves a single resident wave
    with pytest.raises(RuntimeError, match="grouped sweep"):
        srv.submit_dense(dense_int_W(64, 2), 16, 1, sched(1.0, 3), BETA)
    srv = rocm.server()
    N = 11777  # padded to 12032 sites, over the LDS-resident limit of the sweep
    with pytest.raises(RuntimeError, match="grouped sweep"):
        srv.submit_dense(np.zeros((N, N), np.float32), 16, 1, sched(1.0, 3), BETA)
    assert srv.status()["pending"] == 0


def test_failed_job_leaves_the_others_running(rocm):
    """A dense job whose sweep gives up (a spin limit of one blocked
    iteration, forced from the host; with m = 2 both neighbours are one slice
    and its slices wait on each other often) fails with a recorded error,
    its memory is returned, and the jobs stepped beside it -- a dense job in
    the same launch and a bipartite job -- still equal the same jobs run
    alone."""
    js = jobs()
    good_dense = js[0]  # N = 64, m = 16, 30 steps, integer W
    good_bip = js[6]    # bipartite 100 x 80, m = 6
    srv = rocm.server(dense_field_precision="bf16x2", bipartite_field_precision="bf16x2")
    a = submit(srv, good_dense)
    b = submit(srv, good_bip)
    srv.tick()
    srv.set_preferences(dense_spin_limit=1)
    bad = srv.submit_dense(dense_int_W(128, 5), 2, 77, sched(5.0, 30), BETA)
    while srv.tick() > 0:
        pass
    done = set(srv.take_done())
    assert done == {a, b, bad}
    err = srv.error(bad)
    assert err is not None and "spin limit" in err
    with pytest.raises(RuntimeError, match="failed"):
        srv.result(bad)
    assert srv.error(a) is None and srv.error(b) is None
    assert srv.status()["bytes_in_use"] == 0
    for jid, job in ((a, good_dense), (b, good_bip)):
        kind, data, m, seed, G = job
        E, xs = alone(rocm, kind, data, m, seed, G)
        r = srv.result(jid)
        assert np.array_equal(r["E"], E), kind
