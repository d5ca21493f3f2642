"""CPU pinning of the oracle (oracle/, the C restatement of sqaod.cpu).

Each check is chosen so that a plausible restatement mistake fails it:
golden vectors from the reference's own MT19937 compiled from source, the
textbook known answer, an independent MT19937 (numpy's legacy generator),
closed forms and identities from the reference tests, brute force, and a
consistency law between two independent reference formulas (the annealer's
Metropolis dE and getSystemE's system energy).
"""
import json
import os

import numpy as np
import pytest

from problems import dense_int_W, dense_graph_8x8, mat_symmetric_cycle, dense_graph_random

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mt19937_ref.json")


# ---------------------------------------------------------------- MT19937
def test_mt19937_textbook_known_answer(oracle):
    # [rand.predef]: the 10000th output of the default-seeded mt19937 is 4123659995
    assert oracle.mt_stream(5489, 10000)[-1] == 4123659995


def test_mt19937_matches_numpy_legacy_generator(oracle):
    # numpy's RandomState(int) seeds with init_genrand; raw outputs are the tempered words
    for seed in (0, 1, 7, 1133557, 4294967295):
        rs = np.random.RandomState(seed)
        ref = rs._bit_generator.random_raw(3000).astype(np.uint32)
        assert np.array_equal(oracle.mt_stream(seed, 3000), ref)


def test_mt19937_matches_reference_golden(oracle):
    """Golden vectors written by the reference's own Random.cpp (tests/golden/make_golden.py)."""
    g = json.load(open(GOLDEN))
    for case in g["cases"]:
        s = case["seed"]
        assert oracle.mt_stream(s, 24).tolist() == case["u32_first24"]
        assert oracle.mt_stream(s, 8, skip=1000).tolist() == case["u32_skip1000_8"]
        w = oracle.mt_stream(s, 16).astype(np.uint64)
        # randomf32 (Random.cpp:157-161): float(u32) * 2^-32, round-to-nearest-even
        f32 = (w[:8].astype(np.float32) * np.float32(2.0 ** -32)).view(np.uint32)
        assert f32.tolist() == case["f32_first8_bits"]
        # randomf64 (:165-168): ((a>>5)*2^26 + (b>>6)) * 2^-53, a drawn first
        a, b = w[0::2] >> np.uint64(5), w[1::2] >> np.uint64(6)
        f64 = ((a.astype(np.float64) * 67108864.0 + b.astype(np.float64)) * (1.0 / 9007199254740992.0))
        assert [int(v) for v in f64.view(np.uint64)] == case["f64_first8_bits"]
        for n, vals in case["randint_first12"].items():
            assert (oracle.mt_stream(s, 12) % np.uint32(int(n))).tolist() == vals


def test_mt19937_reference_build_if_present(oracle):
    R = oracle.ref_random()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent on this machine)")
    import ctypes
    out = np.empty(5000, np.uint32)
    R.sqref_mt_stream(99, 777, out.ctypes.data_as(ctypes.c_void_p), 5000)
    assert np.array_equal(out, oracle.mt_stream(99, 5000, skip=777))


# ---------------------------------------------------------------- dot, formulas
def test_dot_exact_on_integer_data(oracle):
    rng = np.random.default_rng(3)
    for n in (1, 15, 16, 17, 100, 4096):
        a = rng.integers(-8, 9, n).astype(np.float32) * 0.25
        b = rng.choice([-1.0, 1.0], n).astype(np.float32)
        assert oracle.dot(a, b) == float(np.dot(a.astype(np.float64), b.astype(np.float64)))


def test_hamiltonian_identity_and_energy_closed_form(oracle):
    """test_dense_graph_formulas.py:53-59: -sum(h) + sum(J) + c = 0, and the
    Ising energy of q equals x^T W x with x = (q + 1)/2 on all 2^8 states
    (test_dense_graph_annealer.py:98-120)."""
    for W in (dense_graph_random(8, 5), dense_int_W(8, 11).astype(np.float64), mat_symmetric_cycle(8)):
        h, J, c = oracle.dense_hamiltonian(W)
        assert np.all(np.diag(J) == 0)
        assert abs(-np.sum(h) + np.sum(J) + c) < 1e-9
        xs = ((np.arange(256)[:, None] >> np.arange(7, -1, -1)[None, :]) & 1).astype(np.int8)
        Ex = oracle.dense_energy_x(W, xs)
        Eq = oracle.dense_energy_q(h, J, c, 2 * xs - 1)
        assert np.allclose(Ex, Eq, atol=1e-9)
        assert np.allclose(Ex, np.einsum("ki,ij,kj->k", xs, W, xs), atol=1e-9)


def test_energy_known_answers(oracle):
    """test_dense_graph_annealer.py:85-96, :154-165: W = ones(8), q = +1 -> E = 64; q = -1 -> E = 0."""
    W = np.ones((8, 8))
    h, J, c = oracle.dense_hamiltonian(W)
    assert oracle.dense_energy_q(h, J, c, np.ones((1, 8), np.int8))[0] == 64.0
    assert oracle.dense_energy_q(h, J, c, -np.ones((1, 8), np.int8))[0] == 0.0


# ---------------------------------------------------------------- brute-force search
@pytest.mark.parametrize("opt,sign,Eexp,xbit", [(0, 1, 0.0, 0), (0, -1, -64.0, 1), (1, 1, 64.0, 1), (1, -1, 0.0, 0)])
def test_bf_known_answers(oracle, opt, sign, Eexp, xbit):
    """test_dense_graph_bf_searcher.py:82-97 (+-ones(8), minimize/maximize)."""
    E, xs = oracle.dense_bf_search(sign * np.ones((8, 8)), optimize=opt)
    assert E == Eexp
    assert len(xs) == 1 and np.all(oracle.unpack_bits(xs, 8) == xbit)


def test_bf_matches_direct_enumeration(oracle):
    rng = np.random.default_rng(5)
    for N in (3, 6, 9, 12):
        W = dense_int_W(N, 100 + N, lo=-3, hi=3).astype(np.float64)
        xs_all = ((np.arange(1 << N)[:, None] >> np.arange(N - 1, -1, -1)[None, :]) & 1).astype(np.float64)
        E_all = np.einsum("ki,ij,kj->k", xs_all, W, xs_all)
        Emin = E_all.min()
        ties = np.nonzero(E_all == Emin)[0][:1024]
        E, xs = oracle.dense_bf_search(W, dtype=np.float32)
        assert E == Emin
        assert xs.tolist() == ties.tolist()


def test_bf_tie_list_truncated_to_tile_size(oracle):
    """W = 0: every x ties; the list keeps the tile_size smallest x in order
    (CPUDenseGraphBFSearcher.cpp:104-139)."""
    E, xs = oracle.dense_bf_search(np.zeros((12, 12)), tile_size=256)
    assert E == 0.0 and xs.tolist() == list(range(256))


def test_bf_8x8_example(oracle):
    """example_problems.py:4-14: E(k bits set) = 4k^2 - 36k, minimum -80 at k = 4 and 5."""
    E, xs = oracle.dense_bf_search(dense_graph_8x8())
    bits = oracle.unpack_bits(xs, 8).sum(axis=1)
    assert E == -80.0 and set(bits.tolist()) == {4, 5} and len(xs) == 70 + 56


# ---------------------------------------------------------------- annealer laws
def _trotter_action(an, G, beta):
    """Suzuki-Trotter action S = mean_y E_y + (c/2) sum_y q_y . q_(y+1),
    c = ln(tanh(G beta / m)) / beta (< 0: ferromagnetic between slices)."""
    q = an.get_q().astype(np.int64)
    m = q.shape[0]
    spin_dot = int(np.sum(q * np.roll(q, -1, axis=0)))
    coef = np.log(np.tanh(G * beta / m)) / beta
    return float(np.mean(an.calculate_E())) + 0.5 * coef * spin_dot, spin_dot


def test_zero_temperature_sqa_never_raises_trotter_action(oracle):
    """The Metropolis dE of tryFlip (CPUDenseGraphAnnealer.cpp:250-265) is the
    change of the Suzuki-Trotter action: classical part (2/m) q (h + 2 J.q) =
    dE_classical / m, coupling part -q (q_up + q_down) c.  With beta -> inf
    no accepted flip raises the action."""
    N, m = 24, 6
    W = dense_int_W(N, 77, lo=-4, hi=4)
    an = oracle.DenseGraphAnnealer(np.float64, 1)
    an.seed(5)
    an.set_qubo(W.astype(np.float64))
    an.set_m(m)
    an.prepare()
    an.randomize_spin()
    G, beta = 2.0, 1e6
    prev, _ = _trotter_action(an, G, beta)
    flips = 0
    for _ in range(12):
        an.anneal_one_step(G, beta)
        flips += an.last_flips()
        cur, _ = _trotter_action(an, G, beta)
        assert cur <= prev + 1e-9
        prev = cur
    assert flips > 0


def test_trotter_coupling_aligns_slices(oracle):
    """W = 0 leaves only the inter-slice term: with G beta / m = 1e-6 the
    coupling c = ln tanh(1e-6) is strongly ferromagnetic and an uphill flip has
    probability tanh(1e-6)^2, so the slice overlap sum_y q_y . q_(y+1) never
    decreases and the action never rises.  A sign error in the coupling term
    of dE would anti-align the slices instead."""
    N, m = 16, 6
    an = oracle.DenseGraphAnnealer(np.float64, 1)
    an.seed(8)
    an.set_qubo(np.zeros((N, N)))
    an.set_m(m)
    an.prepare()
    an.randomize_spin()
    beta = 1.0
    G = 1e-6 * m / beta
    prev_S, prev_dot = _trotter_action(an, G, beta)
    for _ in range(20):
        an.anneal_one_step(G, beta)
        S, dot = _trotter_action(an, G, beta)
        assert dot >= prev_dot and S <= prev_S + 1e-9
        prev_S, prev_dot = S, dot
    assert prev_dot > 0


def _action_all(W, q, coef):
    """Trotter action for spins q (m x N) of QUBO W (minimize), in float64."""
    x = (q + 1) // 2
    E = np.einsum("yi,ij,yj->y", x, W, x)
    dot = np.sum(q * np.roll(q, -1, axis=0))
    return E.mean() + 0.5 * coef * dot


def test_sa_zero_temperature_reaches_local_minimum(oracle):
    """SA at kT -> 0 ends in a state no single flip improves: pins the scale
    of the SA dE = 2 q (h + 2 J.q) (CPUDenseGraphAnnealer.cpp:348) against the
    true energy change."""
    for seed in (1, 2, 3):
        N = 20
        W = dense_int_W(N, 300 + seed, lo=-4, hi=4).astype(np.float64)
        an = oracle.DenseGraphAnnealer(np.float64, 1)
        an.seed(seed)
        an.set_qubo(W)
        an.set_m(1)
        an.prepare()
        an.randomize_spin()
        for _ in range(40):
            an.anneal_one_step(1e-9, 0.0)
        q = an.get_q().astype(np.int64)[0]
        x = (q + 1) // 2
        E0 = x @ W @ x
        for i in range(N):
            x1 = x.copy()
            x1[i] ^= 1
            assert x1 @ W @ x1 >= E0


def test_sqa_zero_temperature_reaches_local_minimum_of_trotter_action(oracle):
    """With the coupling chosen comparable to the classical term
    (c = -3/16, m = 4: classical dE in multiples of 1/4, coupling 0 or +-3/8)
    and beta = 1000 (every non-zero |dE| >= 1/8), the final state is a local
    minimum of the Trotter action over all single flips: pins the relative
    scale of 2/m, the factor 2 on J.q, the neighbour slices and the coupling
    term of tryFlip (:258-261)."""
    for seed in (4, 5, 6, 7, 8, 9):
        N, m, beta = 16, 4, 1000.0
        coef = -3.0 / 16
        # coef = ln(tanh(G beta / m)) / beta  =>  G = atanh(exp(coef * beta)) * m / beta
        G = np.arctanh(np.exp(coef * beta)) * m / beta
        W = dense_int_W(N, 400 + seed, lo=-3, hi=3).astype(np.float64)
        an = oracle.DenseGraphAnnealer(np.float64, 1)
        an.seed(seed)
        an.set_qubo(W)
        an.set_m(m)
        an.prepare()
        an.randomize_spin()
        for _ in range(60):
            an.anneal_one_step(G, beta)
        c_eff = np.log(np.tanh(G * beta / m)) / beta
        assert abs(c_eff - coef) < 1e-9
        q = an.get_q().astype(np.int64)
        S0 = _action_all(W, q, coef)
        for y in range(m):
            for i in range(N):
                q1 = q.copy()
                q1[y, i] = -q1[y, i]
                assert _action_all(W, q1, coef) >= S0 - 1e-9


def test_system_energy_restates_reference_formula(oracle):
    """getSystemE (:230-247) = mean E - spinDotSum * (0.5/beta) ln tanh(G beta/m)
    (note the opposite sign of the coupling to the action the dynamics lower)."""
    N, m = 12, 4
    an = oracle.DenseGraphAnnealer(np.float64, 1)
    an.seed(1)
    an.set_qubo(dense_int_W(N, 3).astype(np.float64))
    an.set_m(m)
    an.prepare()
    an.randomize_spin()
    G, beta = 0.7, 3.0
    q = an.get_q().astype(np.int64)
    dot = int(np.sum(q * np.roll(q, -1, axis=0)))
    expect = float(np.mean(an.calculate_E())) - dot * (0.5 / beta) * np.log(np.tanh(G * beta / m))
    assert abs(an.get_system_E(G, beta) - expect) < 1e-9


def test_zero_temperature_sa_never_raises_energy(oracle):
    """SA dE (:341-352) is the energy change: at kT -> 0 energies never rise."""
    N = 32
    W = dense_int_W(N, 78, lo=-4, hi=4)
    an = oracle.DenseGraphAnnealer(np.float32, 1)
    an.seed(6)
    an.set_qubo(W)
    an.set_m(1)
    assert an.prepare() == oracle.ALGO_SA_NAIVE
    an.randomize_spin()
    prev = an.calculate_E()[0]
    for _ in range(10):
        an.anneal_one_step(1e-7, 0.0)
        cur = an.calculate_E()[0]
        assert cur <= prev
        prev = cur


def test_trivial_problems_reach_zero(oracle):
    """test_dense_graph_annealer.py:180-260: W = ones(10), m = 4, G 5 -> 0.02."""
    for algo in (oracle.ALGO_COLORING, oracle.ALGO_NAIVE):
        an = oracle.DenseGraphAnnealer(np.float64, 1)
        an.seed(0)
        an.set_qubo(np.ones((10, 10)))
        an.set_m(4)
        an.set_algorithm(algo)
        an.prepare()
        an.randomize_spin()
        G = 5.0
        while G > 0.02:
            an.anneal_one_step(G, 1 / 0.03)
            G *= 0.95
        assert an.calculate_E().min() == 0.0


def test_randomize_spin_consumes_serial_stream(oracle):
    """randomizeSpin draws randInt(2) row-major from worker 0 (:152-161)."""
    N, m = 10, 3
    an = oracle.DenseGraphAnnealer(np.float32, 1)
    an.seed(42)
    an.set_qubo(np.ones((N, N), np.float32))
    an.set_m(m)
    an.prepare()
    an.randomize_spin()
    w = oracle.mt_stream(42, m * N)
    assert np.array_equal(an.get_q().reshape(-1), np.where(w % 2 == 1, 1, -1))


def test_parallel_coloring_is_deterministic(oracle):
    """The OpenMP coloring variant (:303-339) is deterministic for a fixed
    worker count; with one worker it is the serial order."""
    N, m = 40, 8
    W = dense_int_W(N, 9)
    qs = []
    for _ in range(2):
        an = oracle.DenseGraphAnnealer(np.float32, 4)
        an.seed(3)
        an.set_qubo(W)
        an.set_m(m)
        an.prepare()
        an.randomize_spin()
        for G in (5.0, 3.0, 1.0):
            an.anneal_one_step(G, 10.0)
        qs.append(an.get_q())
    assert np.array_equal(qs[0], qs[1])


def test_host_expf_matches_restatement_on_acceptance_domain(oracle):
    """sq_expf (the device exp) equals the host expf on every float in
    (-8, -0] (bit patterns 0x80000000..0xc1000000); the full 2^32 comparison
    of the device exp runs on the GPU box."""
    import subprocess
    exe = os.path.join(os.path.dirname(oracle.LIB_PATH), "check_expf")
    r = subprocess.run([exe, "0x80000000", "0xc1000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout


def test_exp2_table_regenerates():
    """The 2^(i/32) table of sq_expf.h equals the correctly rounded doubles
    recomputed here with 80-digit decimal arithmetic."""
    import re
    import struct
    from decimal import Decimal, getcontext
    getcontext().prec = 80
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "sqaod_amd", "csrc", "sq_expf.h")).read()
    tab = [int(v, 16) for v in re.findall(r"0x([0-9a-f]{16})ull", hdr)[:32]]
    ln2 = Decimal(2).ln()
    for i in range(32):
        d = float(str((ln2 * i / 32).exp()))
        u = struct.unpack("<Q", struct.pack("<d", d))[0]
        assert tab[i] == (u - (i << 47)) & 0xFFFFFFFFFFFFFFFF


# ---------------------------------------------------------------- bipartite
def _bg_int_problem(N0, N1, seed, lo=-3, hi=3):
    rng = np.random.default_rng(seed)
    b0 = rng.integers(lo, hi + 1, N0).astype(np.float64)
    b1 = rng.integers(lo, hi + 1, N1).astype(np.float64)
    W = rng.integers(lo, hi + 1, (N1, N0)).astype(np.float64)
    return b0, b1, W


def test_bipartite_hamiltonian_reproduces_qubo_energy_on_all_states(oracle):
    """BGFuncs::calculateHamiltonian (SharedFormulas.cpp:147-167) maps the QUBO
    E(x0, x1) = b0.x0 + b1.x1 + x1^T W x0 onto the Ising form
    -(h0.q0 + h1.q1 + q1^T J q0) - c with q = 2x - 1 (the identity
    test_bipartite_graph_annealer.py:98-120 checks through the annealer):
    every (x0, x1) of a 6 x 5 problem, so a dropped term, a wrong sign or a
    transposed W fails."""
    N0, N1 = 6, 5
    b0, b1, W = _bg_int_problem(N0, N1, 11)
    h0, h1, J, c = oracle.bipartite_hamiltonian(b0, b1, W)
    x0 = ((np.arange(1 << N0)[:, None] >> np.arange(N0 - 1, -1, -1)[None, :]) & 1).astype(np.float64)
    x1 = ((np.arange(1 << N1)[:, None] >> np.arange(N1 - 1, -1, -1)[None, :]) & 1).astype(np.float64)
    Eq = (x0 @ b0)[:, None] + (x1 @ b1)[None, :] + x0 @ W.T @ x1.T
    q0, q1 = 2 * x0 - 1, 2 * x1 - 1
    Ei = -((q0 @ h0)[:, None] + (q1 @ h1)[None, :] + q0 @ J.T @ q1.T) - c
    assert np.array_equal(Eq, Ei)


def test_bipartite_annealer_energy_known_answers(oracle):
    """test_bipartite_graph_annealer.py:88-104, 164-177: q = -1 gives E = 0;
    b = 1, W = 1, q = +1 gives E = N0 N1 + N0 + N1; with the weights negated
    and maximize, E is that problem's own energy (the sign is restored)."""
    N0, N1 = 7, 4
    for opt, sgn in ((0, 1.0), (1, -1.0)):
        an = oracle.BipartiteGraphAnnealer(np.float64)
        an.seed(1)
        an.set_qubo(sgn * np.ones(N0), sgn * np.ones(N1), sgn * np.ones((N1, N0)), optimize=opt)
        an.set_m(1)
        an.prepare()
        an.set_q(np.ones((1, N0), np.int8), np.ones((1, N1), np.int8))
        assert an.calculate_E().tolist() == [sgn * (N0 * N1 + N0 + N1)]
        an.set_q(-np.ones((1, N0), np.int8), -np.ones((1, N1), np.int8))
        assert an.calculate_E().tolist() == [0.0]


def test_bipartite_randomize_spin_consumes_serial_stream(oracle):
    """randomizeSpin with one worker (CPUBipartiteGraphAnnealer.cpp:172-192):
    per slice, N0 draws for the q0 row, then N1 for the q1 row."""
    N0, N1, m = 6, 5, 3
    an = oracle.BipartiteGraphAnnealer(np.float32)
    an.seed(42)
    an.set_qubo(np.ones(N0, np.float32), np.ones(N1, np.float32), np.ones((N1, N0), np.float32))
    an.set_m(m)
    an.prepare()
    an.randomize_spin()
    q0, q1 = an.get_q()
    w = np.where(oracle.mt_stream(42, m * (N0 + N1)) % 2 == 1, 1, -1).reshape(m, N0 + N1)
    assert np.array_equal(q0, w[:, :N0]) and np.array_equal(q1, w[:, N0:])


def _bg_action(h0, h1, J, c, q0, q1, coef):
    """Suzuki-Trotter action of the bipartite model: mean_y E_y +
    (coef/2) sum_y (q0_y . q0_(y+1) + q1_y . q1_(y+1))."""
    E = -(q0 @ h0 + q1 @ h1 + np.einsum("yi,ij,yj->y", q1, J, q0)) - c
    dot = np.sum(q0 * np.roll(q0, -1, axis=0)) + np.sum(q1 * np.roll(q1, -1, axis=0))
    return E.mean() + 0.5 * coef * dot


@pytest.mark.parametrize("m", [4, 5])
def test_bipartite_sqa_zero_temperature_reaches_local_minimum_of_trotter_action(oracle, m):
    """tryFlip of the bipartite half-steps (:331-344: dE = (2/m) q (h + dEmat)
    - q (q_prev + q_next) coef, no factor 2 on dEmat) is the change of the
    Trotter action; at beta = 1000 with coef = -3/16 every non-zero |dE| is at
    least 1/10, so the final state is a local minimum over every single flip
    of both parts.  Pins 2/m, the missing factor 2, J against J^T, the
    neighbour slices (m odd: the ring closes between two even slices) and the
    sign of the coupling."""
    for seed in (3, 4, 5):
        N0, N1, beta, coef = 9, 7, 1000.0, -3.0 / 16
        G = np.arctanh(np.exp(coef * beta)) * m / beta
        b0, b1, W = _bg_int_problem(N0, N1, 50 + seed)
        an = oracle.BipartiteGraphAnnealer(np.float64)
        an.seed(seed)
        an.set_qubo(b0, b1, W)
        an.set_m(m)
        an.prepare()
        an.randomize_spin()
        for _ in range(60):
            an.anneal_one_step(G, beta)
        h0, h1, J, c = an.get_hamiltonian()
        q0, q1 = (v.astype(np.float64) for v in an.get_q())
        S0 = _bg_action(h0, h1, J, c, q0, q1, coef)
        for part in (0, 1):
            q = (q0, q1)[part]
            for y in range(m):
                for i in range(q.shape[1]):
                    qf = q.copy()
                    qf[y, i] = -qf[y, i]
                    args = (qf, q1) if part == 0 else (q0, qf)
                    assert _bg_action(h0, h1, J, c, *args, coef) >= S0 - 1e-9


def test_bipartite_sa_zero_temperature_reaches_local_minimum(oracle):
    """SA coloring (:485-496, dE = 2 q (h + dEmat)) at kT -> 0 ends in a state
    no single flip improves, slice by slice."""
    for seed in (1, 2):
        N0, N1, m = 10, 8, 3
        b0, b1, W = _bg_int_problem(N0, N1, 70 + seed)
        an = oracle.BipartiteGraphAnnealer(np.float64)
        an.seed(seed)
        an.set_qubo(b0, b1, W)
        an.set_m(m)
        an.set_algorithm(7)
        an.prepare()
        an.randomize_spin()
        for _ in range(30):
            an.anneal_one_step(1e-9, 0.0)
        h0, h1, J, c = an.get_hamiltonian()
        q0, q1 = (v.astype(np.float64) for v in an.get_q())
        E = lambda a, b: -(a @ h0 + b @ h1 + b @ J @ a) - c
        for y in range(m):
            E0 = E(q0[y], q1[y])
            for i in range(N0):
                a = q0[y].copy()
                a[i] = -a[i]
                assert E(a, q1[y]) >= E0
            for i in range(N1):
                b = q1[y].copy()
                b[i] = -b[i]
                assert E(q0[y], b) >= E0


def test_bipartite_trivial_problems_reach_zero(oracle):
    """test_bipartite_graph_annealer.py:196-266: b = 1, W = 1 anneals to E = 0
    (minimize; maximize with the signs flipped), coloring m = 6 and SA m = 1."""
    N0, N1 = 10, 8
    for algo, m, (Gi, Gf, beta) in ((3, 6, (5.0, 0.02, 1 / 0.03)), (7, 1, (10.0, 0.02, 1.0))):
        for opt, sgn in ((0, 1.0), (1, -1.0)):
            an = oracle.BipartiteGraphAnnealer(np.float32)
            an.seed(12)
            an.set_qubo(sgn * np.ones(N0), sgn * np.ones(N1), sgn * np.ones((N1, N0)), optimize=opt)
            an.set_m(m)
            an.set_algorithm(algo)
            an.prepare()
            an.randomize_spin()
            G, tau = Gi, np.exp(np.log(Gf / Gi) / 100)
            for _ in range(100):
                an.anneal_one_step(G, beta)
                G *= tau
            E = an.calculate_E()
            assert (E.min() if opt == 0 else E.max()) == 0.0


# ---------------------------------------------------------------- detailed balance
def _state_histogram(an, get_states, steps, G, beta):
    counts = {}
    for _ in range(steps):
        an.anneal_one_step(G, beta)
        key = get_states().tobytes()
        counts[key] = counts.get(key, 0) + 1
    return counts


def _total_variation(counts, states, logw):
    p = np.exp(logw - logw.max())
    p /= p.sum()
    total = sum(counts.values())
    emp = np.array([counts.get(s.astype(np.int8).tobytes(), 0) / total for s in states])
    return 0.5 * np.abs(emp - p).sum()


@pytest.mark.parametrize("sa", [False, True])
def test_bipartite_stationary_distribution_is_boltzmann(oracle, sa):
    """Every single-spin Metropolis update of the half-steps leaves the
    Boltzmann distribution of its energy function invariant, so the chain's
    state distribution converges to exp(-beta S) (coloring; S the Trotter
    action, (1/m) sum_y E_y + (coef/2) sum_y q_y.q_(y+1) over both parts) or
    exp(-E/kT) per slice (SA).  On a 2 x 1 problem with m = 3 (512 states) the
    empirical distribution over 40000 steps is within total variation 0.08 of
    the exact one (measured 0.02); a wrong dE scale (2/m, the SA factor 2, a
    factor 2 on dEmat), a wrong neighbour slice or a wrong coupling sign moves
    it by 0.15 or more."""
    import itertools
    b0, b1, W = np.array([0.5, -1.0]), np.array([0.75]), np.array([[1.5, -1.0]])
    N0, N1, m = 2, 1, 3
    an = oracle.BipartiteGraphAnnealer(np.float64)
    an.seed(7)
    an.set_qubo(b0, b1, W)
    an.set_m(m)
    if sa:
        an.set_algorithm(7)
    an.prepare()
    an.randomize_spin()
    h0, h1, J, c = an.get_hamiltonian()
    G, beta = 0.8, 1.0
    states = np.array(list(itertools.product([-1, 1], repeat=(N0 + N1) * m)), np.float64)
    q0 = states[:, :N0 * m].reshape(-1, m, N0)
    q1 = states[:, N0 * m:].reshape(-1, m, N1)
    E = -(q0 @ h0 + q1 @ h1 + np.einsum("syi,ij,syj->sy", q1, J, q0)) - c
    if sa:
        logw = -E.sum(axis=1) / G  # kT = G, slices independent
    else:
        coef = np.log(np.tanh(G * beta / m)) / beta
        dot = np.sum(q0 * np.roll(q0, -1, axis=1), axis=(1, 2)) + np.sum(q1 * np.roll(q1, -1, axis=1), axis=(1, 2))
        logw = -beta * (E.mean(axis=1) + 0.5 * coef * dot)

    def get_states():
        a, b = an.get_q()
        return np.concatenate([a.reshape(-1), b.reshape(-1)]).astype(np.int8)

    counts = _state_histogram(an, get_states, 40000, G, beta)
    assert _total_variation(counts, states, logw) < 0.08


@pytest.mark.parametrize("N,m,G", [(5, 1, 0.7), (3, 3, 0.7), (1, 6, 0.3)])
def test_dense_stationary_distribution_is_boltzmann(oracle, N, m, G):
    """The same law for the dense annealer (random-site Metropolis of tryFlip,
    :250-265, and SA for m = 1, :341-371), 40000 steps, total variation below
    0.08 (measured 0.015-0.025).  N = 1, m = 6 with G = 0.3 is a ring of six
    strongly coupled slices of one site: a wrong wrap-around neighbour of slice
    0 or m-1 moves the distribution by more than 0.9."""
    import itertools
    rng = np.random.default_rng(31)
    U = rng.integers(-2, 3, (N, N)).astype(np.float64) * 0.5
    W = np.triu(U) + np.triu(U, 1).T
    an = oracle.DenseGraphAnnealer(np.float64, 1)
    an.seed(9)
    an.set_qubo(W)
    an.set_m(m)
    an.prepare()
    an.randomize_spin()
    beta = 1.0
    states = np.array(list(itertools.product([-1, 1], repeat=N * m)), np.float64)
    x = (states.reshape(-1, m, N) + 1) / 2
    E = np.einsum("syi,ij,syj->sy", x, W, x)
    if m == 1:
        logw = -E[:, 0] / G  # SA: kT = G
    else:
        coef = np.log(np.tanh(G * beta / m)) / beta
        q = states.reshape(-1, m, N)
        dot = np.sum(q * np.roll(q, -1, axis=1), axis=(1, 2))
        logw = -beta * (E.mean(axis=1) + 0.5 * coef * dot)
    counts = _state_histogram(an, lambda: an.get_q().reshape(-1).astype(np.int8), 40000, G, beta)
    assert _total_variation(counts, states, logw) < 0.08


# ---------------- bipartite brute-force search ----------------

def _bg_all_pairs(b0, b1, W):
    """E[x0, x1] for every packed pair (bit N-1-pos = x[pos]), exact."""
    N0, N1 = len(b0), len(b1)
    X0 = ((np.arange(1 << N0)[:, None] >> np.arange(N0 - 1, -1, -1)[None, :]) & 1).astype(np.float64)
    X1 = ((np.arange(1 << N1)[:, None] >> np.arange(N1 - 1, -1, -1)[None, :]) & 1).astype(np.float64)
    return (X0 @ b0)[:, None] + (X1 @ b1)[None, :] + X0 @ W.T @ X1.T


@pytest.mark.parametrize("opt,sign,Eexp,bit", [(0, 1, 0.0, 0), (0, -1, -(64 + 16), 1), (1, 1, 64 + 16, 1),
                                              (1, -1, 0.0, 0)])
def test_bg_bf_known_answers(oracle, opt, sign, Eexp, bit):
    """test_bipartite_graph_bf_searcher.py:62-110: +-ones problems at 8 x 8."""
    N0 = N1 = 8
    E, xs0, xs1 = oracle.bipartite_bf_search(sign * np.ones(N0), sign * np.ones(N1), sign * np.ones((N1, N0)),
                                             optimize=opt)
    assert E == Eexp and len(xs0) == 1
    full = (1 << 8) - 1
    assert xs0[0] == (full if bit else 0) and xs1[0] == (full if bit else 0)


@pytest.mark.parametrize("N0,N1,T0,seed", [(5, 4, 1024, 1), (3, 7, 4, 2), (6, 6, 16, 3), (7, 2, 32, 4)])
def test_bg_bf_matches_brute_force_in_visiting_order(oracle, N0, N1, T0, seed):
    """Integer b0, b1, W in [-2, 2] (many ties): Emin equals the exhaustive
    minimum, and the list is the first N0 + N1 minimisers in the one-worker
    order (x0 tile, x1, x0 within the tile) -- computed here by sorting keys."""
    rng = np.random.default_rng(seed)
    b0 = rng.integers(-2, 3, N0).astype(np.float64)
    b1 = rng.integers(-2, 3, N1).astype(np.float64)
    W = rng.integers(-2, 3, (N1, N0)).astype(np.float64)
    for opt, sgn in ((0, 1.0), (1, -1.0)):
        E = sgn * _bg_all_pairs(b0, b1, W)
        Emin = E.min()
        x0s, x1s = np.nonzero(E == Emin)
        T = min(T0, 1 << N0)
        key = (x0s // T) * (1 << N1) * T + x1s * T + x0s % T
        order = np.argsort(key)[:N0 + N1]
        Eo, xs0, xs1 = oracle.bipartite_bf_search(b0, b1, W, optimize=opt, tile_size0=T0)
        assert Eo == sgn * Emin
        assert np.array_equal(xs0, x0s[order].astype(np.uint64)) and np.array_equal(xs1, x1s[order].astype(np.uint64))


def test_bg_bf_tie_list_capped(oracle):
    """All 2^(N0+N1) pairs tie at E = 0: the list holds the first N0 + N1 pairs:
    x0 = 0..T0-1 of tile 0 under x1 = 0, then x1 = 1 ..."""
    N0, N1 = 3, 4
    E, xs0, xs1 = oracle.bipartite_bf_search(np.zeros(N0), np.zeros(N1), np.zeros((N1, N0)))
    assert E == 0.0 and len(xs0) == N0 + N1
    assert list(xs0) == [0, 1, 2, 3, 4, 5, 6] and list(xs1) == [0] * 7
    E, xs0, xs1 = oracle.bipartite_bf_search(np.zeros(N0), np.zeros(N1), np.zeros((N1, N0)), tile_size0=2)
    assert list(xs0) == [0, 1, 0, 1, 0, 1, 0] and list(xs1) == [0, 0, 1, 1, 2, 2, 3]
