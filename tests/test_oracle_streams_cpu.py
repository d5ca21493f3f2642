"""The oracle's attempt streams and its fast paths (CPU only).

* Philox4x32-10 against the known-answer vectors published with the
  algorithm (Random123's kat_vectors for philox4x32_10).
* The Philox layouts the rocm device uses (attempt arrays, spins) against
  single Philox blocks and Random.cpp's word -> uniform conversions.
* Injected streams: an annealer consuming (x, u) arrays built from the
  MT19937 stream in tryFlip's draw order (x at CPUDenseGraphAnnealer.cpp:255,
  u at :263) equals the annealer drawing them itself, for coloring, SA and
  naive (:268-280, where y is drawn first).
* Local fields: the incremental-field annealers equal the per-attempt dot /
  per-half-step GEMM annealers bit for bit on integer-valued problems, in
  every algorithm and worker count.
"""
import numpy as np
import pytest

from problems import dense_int_W

# Random123 kat_vectors, philox4x32_10: (counter, key) -> output
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, out):
    assert list(oracle.philox4x32_10(ctr, key)) == out


def _u32_to_f32(w):
    return (np.asarray(w, np.uint32).astype(np.float32) * np.float32(2.0 ** -32)).astype(np.float32)


def test_philox_attempt_layout(oracle):
    """Attempt (y, k) of sweep s: block at counter (k, y, s_lo, s_hi) keyed by
    (seed_lo, seed_hi); site = word 0 % N, float u = float(word 1) 2^-32,
    double u = ((w1 >> 5) 2^26 + (w2 >> 6)) 2^-53."""
    seed, sweep, N, m = 0x1234_5678_9abc, (3 << 32) | 7, 37, 5
    X, U = oracle.philox_attempts(seed, sweep, N, m, np.float32)
    Xd, Ud = oracle.philox_attempts(seed, sweep, N, m, np.float64)
    key = [seed & 0xffffffff, seed >> 32]
    for y, k in [(0, 0), (1, 5), (4, 36), (2, 17)]:
        r = oracle.philox4x32_10([k, y, sweep & 0xffffffff, sweep >> 32], key)
        assert X[y, k] == r[0] % N and Xd[y, k] == X[y, k]
        assert U[y, k] == _u32_to_f32(r[1])
        a, b = int(r[1]) >> 5, int(r[2]) >> 6
        assert Ud[y, k] == (a * 67108864.0 + b) * 2.0 ** -53
    assert U.max() <= 1.0 and U.min() >= 0.0


def test_philox_spin_layout(oracle):
    seed, salt, N, m = 99, 12, 11, 3
    q = oracle.philox_spins(seed, salt, N, m)
    for y in range(m):
        for x in range(N):
            r = oracle.philox4x32_10([x // 4, y, salt, 0xffffffff], [seed, 0])
            assert q[y, x] == (1 if r[x % 4] & 1 else -1)


def _dense(oracle, W, m, seed, dtype=np.float32, workers=1, algo=None, local=False, stream=0):
    an = oracle.DenseGraphAnnealer(dtype, workers)
    an.seed(seed)
    an.set_qubo(W)
    an.set_m(m)
    if algo is not None:
        an.set_algorithm(algo)
    an.set_stream(stream)
    an.prepare()
    if local:
        an.set_local_fields(True)
    return an


def _mt_injected_step_arrays(oracle, seed, skip, N, m, dtype, algo):
    """(Y, X, U) of one step from the serial MT19937 stream, in the order the
    reference draws them."""
    per = 2 if dtype == np.float32 else 3
    nU = per - 1
    if algo == oracle.ALGO_NAIVE:
        w = oracle.mt_stream(seed, N * m * (per + 1), skip).reshape(N * m, per + 1).astype(np.uint64)
        Y = (w[:, 0] % m).astype(np.uint32)
        X = (w[:, 1] % N).astype(np.uint32)
        U = _uniforms(w[:, 2:], dtype)
        return Y, X, U, skip + N * m * (per + 1)
    w = oracle.mt_stream(seed, N * m * per, skip).reshape(-1, per).astype(np.uint64)
    X = np.empty((m, N), np.uint32)
    U = np.empty((m, N), dtype)
    pos = 0
    order = [(y, k) for y in range(m) for k in range(N)] if algo == oracle.ALGO_SA_NAIVE else \
        [(y, k) for k in range(N) for off in (0, 1) for y in range(off, m, 2)]
    for y, k in order:
        X[y, k] = w[pos, 0] % N
        U[y, k] = _uniforms(w[pos:pos + 1, 1:1 + nU], dtype)[0]
        pos += 1
    return None, X, U, skip + N * m * per


def _uniforms(words, dtype):
    words = np.asarray(words, np.uint64)
    if dtype == np.float32:
        return _u32_to_f32(words[:, 0])
    a, b = words[:, 0] >> np.uint64(5), words[:, 1] >> np.uint64(6)
    return (a.astype(np.float64) * 67108864.0 + b.astype(np.float64)) * 2.0 ** -53


@pytest.mark.parametrize("algo", ["coloring", "sa", "naive"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_injected_stream_replays_the_mt_stream(oracle, algo, dtype):
    a = {"coloring": oracle.ALGO_COLORING, "sa": oracle.ALGO_SA_NAIVE, "naive": oracle.ALGO_NAIVE}[algo]
    N, m, seed = 24, 6, 77
    W = dense_int_W(N, 5, dtype=dtype)
    ref = _dense(oracle, W, m, seed, dtype, algo=a)
    inj = _dense(oracle, W, m, seed, dtype, algo=a)
    ref.randomize_spin()
    inj.randomize_spin()
    skip = N * m
    for G in (3.0, 1.0, 0.3):
        beta = 2.0 if a != oracle.ALGO_SA_NAIVE else 1.0
        ref.anneal_one_step(G, beta)
        Y, X, U, skip = _mt_injected_step_arrays(oracle, seed, skip, N, m, dtype, a)
        inj.anneal_one_step_injected(G, beta, X, U, Y)
        assert np.array_equal(ref.get_q(), inj.get_q())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_philox_stream_is_the_injected_device_layout(oracle, dtype):
    """STREAM_PHILOX: randomize_spin takes sweep index 0, step s takes s + 1."""
    N, m, seed = 30, 4, 2024
    W = dense_int_W(N, 6, dtype=dtype)
    ph = _dense(oracle, W, m, seed, dtype, stream=oracle.STREAM_PHILOX)
    inj = _dense(oracle, W, m, seed, dtype)
    ph.randomize_spin()
    inj.set_q(oracle.philox_spins(seed, 0, N, m))
    assert np.array_equal(ph.get_q(), inj.get_q())
    for s, G in enumerate((4.0, 2.0, 0.5, 0.1)):
        ph.anneal_one_step(G, 3.0)
        X, U = oracle.philox_attempts(seed, s + 1, N, m, dtype)
        inj.anneal_one_step_injected(G, 3.0, X, U)
        assert np.array_equal(ph.get_q(), inj.get_q())


@pytest.mark.parametrize("algo,workers,stream", [("coloring", 1, 0), ("coloring", 3, 0), ("coloring", 1, 1),
                                                 ("sa", 1, 0), ("naive", 1, 0), ("naive", 1, 1)])
def test_dense_local_fields_equal_the_dot(oracle, algo, workers, stream):
    a = {"coloring": oracle.ALGO_COLORING, "sa": oracle.ALGO_SA_NAIVE, "naive": oracle.ALGO_NAIVE}[algo]
    N, m, seed = 48, 7, 11
    W = dense_int_W(N, 12)
    dot = _dense(oracle, W, m, seed, workers=workers, algo=a, stream=stream)
    loc = _dense(oracle, W, m, seed, workers=workers, algo=a, stream=stream, local=True)
    dot.randomize_spin()
    loc.randomize_spin()
    G, flips = 5.0, 0
    for _ in range(60):
        beta = 0.05 if a != oracle.ALGO_SA_NAIVE else 1.0
        G2 = G if a != oracle.ALGO_SA_NAIVE else 40.0 * G
        dot.anneal_one_step(G2, beta)
        loc.anneal_one_step(G2, beta)
        flips += dot.last_flips()
        assert loc.last_flips() == dot.last_flips()
        G *= 0.93
    assert flips > 100
    assert np.array_equal(dot.get_q(), loc.get_q())
    assert np.array_equal(dot.calculate_E(), loc.calculate_E())


def _bipartite(oracle, N0, N1, m, seed, workers=1, algo=None, local=False, stream=0):
    rng = np.random.default_rng(seed)
    b0 = rng.integers(-8, 9, N0).astype(np.float32)
    b1 = rng.integers(-8, 9, N1).astype(np.float32)
    W = rng.integers(-8, 9, (N1, N0)).astype(np.float32)
    an = oracle.BipartiteGraphAnnealer(np.float32, workers)
    an.seed(seed)
    an.set_qubo(b0, b1, W)
    an.set_m(m)
    if algo is not None:
        an.set_algorithm(algo)
    an.set_stream(stream)
    an.prepare()
    if local:
        an.set_local_fields(True)
    return an


@pytest.mark.parametrize("algo,workers,stream,m", [("coloring", 1, 0, 6), ("coloring", 3, 0, 7),
                                                   ("coloring", 1, 1, 5), ("sa", 1, 1, 4)])
def test_bipartite_local_fields_equal_the_gemm(oracle, algo, workers, stream, m):
    a = oracle.ALGO_COLORING if algo == "coloring" else oracle.ALGO_SA_COLORING
    g = _bipartite(oracle, 40, 33, m, 5, workers, a, False, stream)
    loc = _bipartite(oracle, 40, 33, m, 5, workers, a, True, stream)
    g.randomize_spin()
    loc.randomize_spin()
    G = 5.0
    for _ in range(40):
        G2 = G if a == oracle.ALGO_COLORING else 30.0 * G
        g.anneal_one_step(G2, 0.1)
        loc.anneal_one_step(G2, 0.1)
        G *= 0.9
    q0, q1 = g.get_q()
    l0, l1 = loc.get_q()
    assert np.array_equal(q0, l0) and np.array_equal(q1, l1)


def test_bipartite_workers_change_the_stream_only(oracle):
    """With several workers the bipartite annealer draws per worker
    (CPUBipartiteGraphAnnealer.cpp:377-431): a different trajectory than one
    worker, deterministic, and with one worker the serial order."""
    runs = []
    for workers in (1, 1, 4, 4):
        an = _bipartite(oracle, 20, 24, 6, 9, workers)
        an.randomize_spin()
        for G in (3.0, 1.0):
            an.anneal_one_step(G, 0.5)
        runs.append(np.concatenate([q.ravel() for q in an.get_q()]))
    assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[2], runs[3])
    assert not np.array_equal(runs[0], runs[2])
