"""The C++ interface (include/sqaod_rocm.hpp): tests/cpp/cpp_api_run, built
by __graft_entry__.build(), creates a Device and the solvers through the
sqaod::rocm factories and calls them through the Solver.h-shaped virtual
methods (sqaodc/common/Solver.h:18-252); what they return is checked here
against the oracle, bit for bit on integer problems with the device MT19937
stream."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_bin", "cpp_api_run")


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    path = tmp_path_factory.mktemp("cpp") / "out.json"
    r = subprocess.run([EXE, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with open(path) as f:
        return json.load(f)


def test_dense_anneal_through_the_cpp_factory(oracle, out):
    N, m, algo = out["dense_shape"]
    assert (N, m, algo) == (128, 32, 3) and out["dense_pref_n_trotters"] == [32]
    W = np.array(out["dense_W"], np.float32).reshape(N, N)
    ora = oracle.DenseGraphAnnealer(np.float32, 1)
    ora.seed(1000)
    ora.set_qubo(W)
    ora.set_m(m)
    ora.prepare()
    ora.randomize_spin()
    G = np.array(out["dense_G"], np.float32)
    assert len(G) == 205
    for g in G:
        ora.anneal_one_step(float(g), 50.0)
    q = np.array(out["dense_q"], np.int8).reshape(m, N)
    assert np.array_equal(q, ora.get_q())
    assert np.array_equal(np.array(out["dense_x"], np.int8).reshape(m, N), (q + 1) // 2)
    assert np.array_equal(np.array(out["dense_E"], np.float32), ora.calculate_E())
    sysE = ora.get_system_E(float(G[-1]), 50.0)
    assert out["dense_systemE"][0] == pytest.approx(sysE, rel=1e-5)


def test_bipartite_anneal_through_the_cpp_factory(oracle, out):
    b0 = np.array(out["bip_b0"], np.float32)
    b1 = np.array(out["bip_b1"], np.float32)
    W = np.array(out["bip_W"], np.float32).reshape(len(b1), len(b0))
    ora = oracle.BipartiteGraphAnnealer(np.float32, 1)
    ora.seed(11)
    ora.set_qubo(b0, b1, W)
    ora.set_m(8)
    ora.prepare()
    ora.randomize_spin()
    for g in np.array(out["dense_G"], np.float32)[:60]:
        ora.anneal_one_step(float(g), 50.0)
    r0, r1 = ora.get_q()
    assert np.array_equal(np.array(out["bip_q0"], np.int8).reshape(8, len(b0)), r0)
    assert np.array_equal(np.array(out["bip_q1"], np.int8).reshape(8, len(b1)), r1)
    assert np.array_equal(np.array(out["bip_E"], np.float32), ora.calculate_E())


def test_bf_search_and_formulas_through_the_cpp_factory(oracle, out):
    Wbf = np.array(out["bf_W"], np.float32).reshape(12, 12)
    Emin, xs = oracle.dense_bf_search(Wbf)
    E = np.array(out["bf_E"], np.float32)
    assert len(E) == len(xs) and np.all(E == np.float32(Emin))
    assert np.array_equal(np.array(out["bf_x"], np.int8).reshape(-1, 12), oracle.unpack_bits(xs, 12))
    assert out["bf_x_end"][0] == 2 ** 12
    N = 128
    W = np.array(out["dense_W"], np.float64).reshape(N, N)
    X = np.array(out["formulas_x"], np.float64).reshape(5, N)
    ref = np.einsum("bi,bi->b", X @ W, X)
    assert np.array_equal(np.array(out["formulas_E"], np.float32), ref.astype(np.float32))
    assert out["state_error_thrown"] == [1]
