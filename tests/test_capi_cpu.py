"""The C-ABI library (include/sqaod_rocm.h) loads and exports every entry
point the header declares, and the host-side solver logic (state machine,
preferences, argument errors) behaves like the reference's without a GPU:
every case here raises before any device call."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sqaod_rocm.h")


@pytest.fixture(scope="module")
def cext():
    from sqaod_amd.rocm import _cext
    return _cext


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b((?:sqaod_rocm|sqaodc_rocm)_\w+)\s*\(", txt)))


def test_header_symbols_exported(cext):
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(cext._lib, s), s
    assert set(syms) == set(cext.EXPORTED)


def test_version_probe_like_the_cuda_library(cext):
    """sqaodc_rocm_version(int *version, int *hip_version), the shape of
    sqaodc_cuda_version (sqaodc/cuda/api_cuda.cpp:5-9), read the way
    sqaodpy/sqaod/common/envcheck.py:27-37 reads the CUDA probe."""
    import ctypes
    so = ctypes.CDLL(cext.LIB_PATH)
    assert hasattr(so, "sqaodc_rocm_version")
    ver, hip_ver = ctypes.c_int(0), ctypes.c_int(0)
    so.sqaodc_rocm_version(ctypes.byref(ver), ctypes.byref(hip_ver))
    assert ver.value == 10003
    assert hip_ver.value >= 70000000  # ROCm 7 HIP: major * 10^7 + minor * 10^5 + patch
    assert cext.version() == (ver.value, hip_ver.value)


def test_state_machine_errors_without_gpu(cext):
    """Solver.cpp:96-112 messages; anneal before prepare raises
    (test_dense_graph_annealer.py:390-393)."""
    A = cext.rocm_dg_annealer
    h = A.new(np.float32)
    try:
        with pytest.raises(RuntimeError, match="Bits\\(x or q\\) not initialized"):
            A.anneal_one_step(h, 1.0, 1.0, np.float32)
        with pytest.raises(RuntimeError, match="Problem is not set"):
            A.prepare(h, np.float32)
        with pytest.raises(RuntimeError, match="Problem is not set"):
            A.get_problem_size(h, np.float32)
        with pytest.raises(RuntimeError, match="positive integer"):
            A.set_preferences(h, {"n_trotters": 0}, np.float32)
        with pytest.raises(RuntimeError, match="unknown preference"):
            A.set_preferences(h, {"no_such_pref": 1}, np.float32)
        with pytest.raises(RuntimeError, match="unknown rng"):
            A.set_preferences(h, {"rng": "xorshift"}, np.float32)
        with pytest.raises(RuntimeError, match="not initialized"):
            A.calculate_E(h, np.float32)
    finally:
        A.delete(h, np.float32)


def test_preferences_defaults_and_algorithm_fallback(cext):
    """Default algorithm coloring; naive is offered, as on the CPU annealer
    (CPUDenseGraphAnnealer.cpp:35-50, test_dense_graph_annealer.py:277-290);
    sa_default and sa_coloring map to sa_naive (Solver.cpp:152-168)."""
    A = cext.rocm_dg_annealer
    h = A.new(np.float64)
    try:
        p = A.get_preferences(h, np.float64)
        assert p["algorithm"] == "coloring" and p["device"] == "rocm" and p["precision"] == "double"
        A.set_preferences(h, {"algorithm": "naive"}, np.float64)
        assert A.get_preferences(h, np.float64)["algorithm"] == "naive"
        A.set_preferences(h, {"algorithm": "sa_coloring"}, np.float64)
        assert A.get_preferences(h, np.float64)["algorithm"] == "sa_naive"
        A.set_preferences(h, {"algorithm": "sa_default"}, np.float64)
        assert A.get_preferences(h, np.float64)["algorithm"] == "sa_naive"
        A.set_preferences(h, {"algorithm": "default", "n_trotters": 7}, np.float64)
        p = A.get_preferences(h, np.float64)
        assert p["algorithm"] == "coloring" and p["n_trotters"] == 7
        with pytest.raises(RuntimeError, match="bf16x2"):
            A.set_preferences(h, {"field_precision": "bf16x2"}, np.float64)
    finally:
        A.delete(h, np.float64)


def test_bad_dtype_and_shapes(cext):
    A = cext.rocm_dg_annealer
    with pytest.raises(RuntimeError):
        A.new(np.int32)
    h = A.new(np.float32)
    try:
        with pytest.raises(RuntimeError, match="square"):
            A.set_qubo(h, np.ones((3, 4), np.float32), 0, np.float32)
        with pytest.raises(RuntimeError, match="optimize"):
            A.set_qubo(h, np.ones((3, 3), np.float32), 5, np.float32)
    finally:
        A.delete(h, np.float32)


def test_python_front_end_mirrors_reference_module():
    """sqaod_amd mirrors the sqaod top level and the device package layout."""
    import sqaod_amd as sq
    import sqaod_amd.rocm as rocm
    assert sq.algorithm.coloring == "coloring" and int(sq.minimize) == 0 and int(sq.maximize) == 1
    assert callable(rocm.dense_graph_annealer) and callable(rocm.dense_graph_bf_searcher)
    x = sq.create_bitset_sequence([5], 4)
    assert x.tolist() == [[0, 1, 0, 1]]
    assert sq.bit_to_spin(x).tolist() == [[-1, 1, -1, 1]]
    W = np.triu(np.ones((4, 4)))
    assert np.array_equal(sq.symmetrize(W), (W + W.T) * 0.5)


def test_bf_searcher_host_logic_without_gpu(cext):
    """Dense BF searcher: set_qubo checks (N <= 63, CUDADenseGraphBFSearcher.cpp:
    54-67; symmetric W), tile_size rounding (Solver.cpp:201-209), algorithm
    pinned to brute_force_search, and the state machine -- all before any
    device call."""
    B = cext.rocm_dg_bf_searcher
    h = B.new(np.float64)
    try:
        with pytest.raises(RuntimeError, match="Problem is not set"):
            B.get_problem_size(h, np.float64)
        with pytest.raises(RuntimeError, match="Problem is not set"):
            B.prepare(h, np.float64)
        with pytest.raises(RuntimeError, match="smaller than 64"):
            B.set_qubo(h, np.zeros((64, 64)), 0, np.float64)
        with pytest.raises(RuntimeError, match="not symmetric"):
            B.set_qubo(h, np.triu(np.ones((5, 5))), 0, np.float64)
        B.set_qubo(h, np.ones((63, 63)), 1, np.float64)
        assert B.get_problem_size(h, np.float64) == 63
        with pytest.raises(RuntimeError, match="not prepared"):
            B.search_range(h, np.float64)
        with pytest.raises(RuntimeError, match="not prepared"):
            B.make_solution(h, np.float64)
        prefs = B.get_preferences(h, np.float64)
        assert prefs == {"algorithm": "brute_force_search", "tile_size": 1024, "precision": "double",
                         "device": "rocm"}
        B.set_preferences(h, {"tile_size": 1000, "algorithm": "coloring"}, np.float64)
        assert B.get_preferences(h, np.float64)["tile_size"] == 1024
        B.set_preferences(h, {"tile_size": 1025}, np.float64)
        assert B.get_preferences(h, np.float64)["tile_size"] == 1280
        assert B.get_preferences(h, np.float64)["algorithm"] == "brute_force_search"
        with pytest.raises(RuntimeError, match="positive integer"):
            B.set_preferences(h, {"tile_size": 0}, np.float64)
    finally:
        B.delete(h, np.float64)


def test_cpp_interface_header_compiles(tmp_path):
    """include/sqaod_rocm.hpp: every solver of the Solver.h-shaped interface,
    for float and double, instantiated in full (all virtual methods), and the
    reference's calling idioms (setPreference(name, value), selectAlgorithm,
    deleteInstance) compile without warnings."""
    import subprocess
    src = tmp_path / "t.cpp"
    kinds = ["DenseGraphAnnealerImpl", "BipartiteGraphAnnealerImpl", "DenseGraphBFSearcherImpl",
             "BipartiteGraphBFSearcherImpl", "DenseGraphFormulasImpl", "BipartiteGraphFormulasImpl"]
    body = "".join(f"template class sqaod::rocm::detail::{k}<{t}>;\n" for k in kinds for t in ("float", "double"))
    body += """
void use() {
    namespace sr = sqaod::rocm;
    sr::Device *d = sr::newDevice();
    sr::DenseGraphAnnealer<double> *a = sr::newDenseGraphAnnealer<double>();
    a->setPreference(sr::pnNumTrotters, 8);
    a->setPreference(sr::pnAlgorithm, sr::algoColoring);
    a->selectAlgorithm(sr::algoSANaive);
    sr::BipartiteGraphBFSearcher<float> *b = sr::newBipartiteGraphBFSearcher<float>();
    sr::PackedBitSet x0 = 0, x1 = 0;
    (void)b->searchRange(&x0, &x1);
    sr::deleteInstance(a);
    sr::deleteInstance(b);
    sr::deleteInstance(d);
}
"""
    src.write_text('#include "sqaod_rocm.hpp"\n' + body)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
