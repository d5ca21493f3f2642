"""The bipartite configuration at its full size: N0 = N1 = 1024, m = 64,
b0, b1, W i.i.d. N(0, 1) float32 (default_rng(1133557), drawn in that
order), the schedule G = 5.0 * 0.97^k while G > 0.01 (205 steps), beta = 50.

* Energies: get_E of the rocm bipartite annealer equals a host FP64
  recompute b0.x0 + b1.x1 + x1^T W x0 of the returned x within 1e-5
  relative (BGFuncs::calculate_E, SharedFormulas.cpp:184-202).
* Distribution: the per-seed minimum and mean slice energies over 1000
  annealer seeds (1000 + 136 i, so that no two CPU runs share a worker
  stream) match the CPU annealer's (two-sample Kolmogorov-Smirnov,
  p > 0.01).  The CPU sample is tests/golden/bipartite_n1024_m64_energies.json,
  written by tests/golden/make_bipartite_ks.py with the oracle's OpenMP
  bipartite annealer (8 workers, MT19937 per worker); the device runs its
  default Philox stream, so the two runs share the law and not the stream.
"""
import json
import os

import numpy as np
import pytest

from problems import schedule

pytestmark = pytest.mark.gpu

N0 = N1 = 1024
M, BETA, PROBLEM_SEED = 64, 50.0, 1133557
SEEDS = [1000 + 8 * 17 * i for i in range(1000)]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bipartite_n1024_m64_energies.json")


@pytest.fixture(scope="module")
def rocm():
    import sqaod_amd.rocm as rocm
    return rocm


@pytest.fixture(scope="module")
def problem():
    rng = np.random.default_rng(PROBLEM_SEED)
    b0 = rng.standard_normal(N0).astype(np.float32)
    b1 = rng.standard_normal(N1).astype(np.float32)
    W = rng.standard_normal((N1, N0)).astype(np.float32)
    return b0, b1, W


def host_energy(b0, b1, W, x0, x1):
    x0 = np.asarray(x0, np.float64)
    x1 = np.asarray(x1, np.float64)
    return x0 @ b0.astype(np.float64) + x1 @ b1.astype(np.float64) + np.einsum(
        "yi,yi->y", x1 @ W.astype(np.float64), x0)


@pytest.fixture(scope="module")
def gpu_sample(rocm, problem):
    """(seed, E[m], x0, x1) for the CPU sample's seeds on one annealer object."""
    b0, b1, W = problem
    an = rocm.bipartite_graph_annealer(b0, b1, W, dtype=np.float32, n_trotters=M)
    an.prepare()
    sch = schedule()
    out = []
    for i, s in enumerate(SEEDS):
        an.seed(s)
        an.randomize_spin()
        for G in sch:
            an.anneal_one_step(G, BETA)
        an.make_solution()
        E = an.get_E().astype(np.float64)
        if i < 10:
            xs = an.get_x()
            out.append((s, E, np.stack([p[0] for p in xs]), np.stack([p[1] for p in xs])))
        else:
            out.append((s, E, None, None))
    return out


def test_energies_match_fp64(problem, gpu_sample):
    b0, b1, W = problem
    for s, E, x0, x1 in gpu_sample[:10]:
        Eh = host_energy(b0, b1, W, x0, x1)
        rel = np.max(np.abs(E - Eh) / np.abs(Eh))
        assert rel <= 1e-5, f"seed {s}: {rel:.3g}"


def test_energy_distribution_matches_cpu_annealer(gpu_sample):
    from scipy.stats import ks_2samp
    with open(GOLDEN) as f:
        cpu = json.load(f)
    assert cpu["seeds"] == SEEDS
    gmin = np.array([E.min() for _, E, _, _ in gpu_sample])
    gmean = np.array([E.mean() for _, E, _, _ in gpu_sample])
    p_min = ks_2samp(gmin, np.array(cpu["E_min"])).pvalue
    p_mean = ks_2samp(gmean, np.array(cpu["E_mean"])).pvalue
    assert p_min > 0.01 and p_mean > 0.01, (p_min, p_mean, gmin.mean(), np.mean(cpu["E_min"]))
