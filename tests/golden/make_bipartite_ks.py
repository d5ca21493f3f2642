"""Writes tests/golden/bipartite_n1024_m64_energies.json: the energy sample of
the CPU bipartite annealer on the config-2 problem, for the two-sample
Kolmogorov-Smirnov test of the rocm bipartite annealer
(tests/test_bipartite_n1024_gpu.py).

Problem: N0 = N1 = 1024, m = 64, b0, b1 and W (N1 x N0) i.i.d. N(0, 1) in
float32 from numpy's default_rng(1133557) drawn in that order; schedule
G = 5.0 * 0.97^k while G > 0.01 (205 steps), beta = 50.

Annealer: the oracle's restatement of sqaod.cpu's OpenMP bipartite annealer
(CPUBipartiteGraphAnnealer.cpp:331-373, the coloring half-steps with the
slices of a colour split in static chunks over 8 workers, worker w drawing
from its own MT19937 seeded s + 17 w, CPUBipartiteGraphAnnealer.cpp:25-33)
with the annealer seeds 1000 + 136 i, i = 0 .. 999, one anneal per seed: seed,
randomize_spin, 205 steps, then E of every slice
(CPUBipartiteGraphAnnealer.cpp:215-222).  Each seed contributes min_y E_y and
mean_y E_y.  The seeds are 136 = 8 x 17 apart because worker w of seed s
draws from the generator seeded s + 17 w: with consecutive seeds, seed s's
worker w + 1 replays seed s + 17's worker w, and the energies of seeds 17
apart correlate (0.6 measured), which breaks the independence the KS test
assumes.

About 16 s per seed on one core (40 minutes with 6 worker processes):
    KS_PROCS=6 python tests/golden/make_bipartite_ks.py
"""
import json
import os
import sys
from multiprocessing import Pool

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "bipartite_n1024_m64_energies.json")

N0 = N1 = 1024
M = 64
BETA = 50.0
N_WORKERS = 8
SEEDS = [1000 + 8 * 17 * i for i in range(1000)]  # no two runs share a worker stream
PROBLEM_SEED = 1133557


def problem():
    import numpy as np
    rng = np.random.default_rng(PROBLEM_SEED)
    b0 = rng.standard_normal(N0).astype(np.float32)
    b1 = rng.standard_normal(N1).astype(np.float32)
    W = rng.standard_normal((N1, N0)).astype(np.float32)
    return b0, b1, W


def schedule():
    out, G = [], 5.0
    while G > 0.01:
        out.append(G)
        G *= 0.97
    return out


def anneal(seed):
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    b0, b1, W = problem()
    a = O.BipartiteGraphAnnealer(np.float32, N_WORKERS)
    a.seed(seed)
    a.set_qubo(b0, b1, W)
    a.set_m(M)
    a.prepare()
    a.randomize_spin()
    for G in schedule():
        a.anneal_one_step(G, BETA)
    E = a.calculate_E().astype(np.float64)
    return seed, float(E.min()), float(E.mean())


def main():
    # the 8 workers are emulated in one thread per process (OMP_THREAD_LIMIT
    # also caps the annealer's num_threads(8) teams); seeds run in parallel
    os.environ["OMP_NUM_THREADS"] = "1"
    os.environ["OMP_THREAD_LIMIT"] = "1"
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    O.build()
    with Pool(int(os.environ.get("KS_PROCS", os.cpu_count() or 1))) as pool:
        rows = sorted(pool.map(anneal, SEEDS, chunksize=4))
    out = {
        "problem": f"bipartite N0=N1={N0}, m={M}, b0, b1, W ~ N(0,1) float32 from default_rng({PROBLEM_SEED})",
        "schedule": "G = 5.0 * 0.97^k while G > 0.01 (205 steps), beta = 50",
        "annealer": f"oracle OpenMP bipartite coloring annealer, {N_WORKERS} workers (MT19937 seeded s + 17 w)",
        "seeds": [r[0] for r in rows],
        "E_min": [r[1] for r in rows],
        "E_mean": [r[2] for r in rows],
    }
    with open(OUT, "w") as f:
        json.dump(out, f)
    print(f"wrote {OUT}: {len(rows)} seeds")


if __name__ == "__main__":
    main()
