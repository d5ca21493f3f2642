#!/usr/bin/env python3
"""Dense SQA throughput on one MI355X: spin-flip attempts/s of annealOneStep at
N=4096 spins, m=256 Trotter slices (W symmetric i.i.d. N(0,1), fp32), over the
annealing schedule G = 5.0 * 0.97^k, beta = 50, one sweep (N*m attempts) per
step.  Four forms of the rocm dense annealer are timed on the same instance:

  bf16x2  local fields in LDS, J rows / field GEMM as bf16 hi+lo, fp32 fields (headline)
  fp32    local fields in LDS, fp32 J
  fp8     local fields in LDS, J as e4m3 with a power-of-two scale per row
          (1-byte J rows, fp8 MFMA field GEMM), fp32 fields
  plain   the reference algorithm: a J-row . q-row dot per attempt, fp32

Every form reports the FP64 energies of its final spins and the max
|E_form - E| of the J it anneals against the given J (both FP64).

Each form runs W warm-up steps, re-randomizes the spins, then K timed steps
bracketed by device synchronisation.  The sweep kernel's average duration is
measured with HIP events on the solver's own stream (profile preference).  The
CPU baseline is the C restatement of sqaod.cpu's OpenMP coloring annealer
timed on the host cores for a bounded number of steps.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_MFMA_PEAK_TFLOPS = 2500.0
FP8_MFMA_PEAK_TFLOPS = 5000.0
PROFILES = os.path.join(ROOT, "profiles")
BETA = 50.0


def traffic_profile_path(N, m, form, steps, warmup):
    """The committed rocprofv3 PMC summary of exactly this bench window."""
    return os.path.join(PROFILES, f"dense_sweep_{form}_n{N}_m{m}_steps0-{steps - 1}_warmup{warmup}.md")


def measured_traffic(N, m, form, steps, warmup):
    """HBM bytes per sweep launch, averaged over the timed launches, from the
    committed rocprofv3 PMC summary of this exact workload and window
    (FETCH_SIZE and WRITE_SIZE runs, calibrated against a device copy of known
    size); None when no such profile exists."""
    path = traffic_profile_path(N, m, form, steps, warmup)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("sweep_hbm_bytes_per_launch:"):
            return int(line.split(":", 1)[1])
    return None


def sweep_bytes_per_launch(N, m, form, flips_per_launch):
    """Bytes the local-field sweep must move in one launch (one annealOneStep,
    N m attempts): per attempt the site and uniform it reads (4 + 4 B), the
    decision word it publishes (4 B) and the two neighbour slices' reads of
    it (2 x 4 B); per slice its fields in and out and h (3 x 4 Np B) and its
    spin bits in and out plus the two neighbour mirrors (4 x Np / 8 B); per
    accepted flip one J row (4 Np B as fp32 / bf16 hi+lo sums, Np + 4 B as
    e4m3 bytes and the row scale)."""
    Np = (N + 255) // 256 * 256
    row = (Np + 1023) // 1024 * 1024 + 4 if form == "fp8" else 4 * Np
    return N * m * 20 + m * (12 * Np + Np // 2) + flips_per_launch * row


def schedule_G(k):
    """G of step k: 5.0 * 0.97^k while G > 0.01 (205 steps), then it repeats."""
    n = 205
    return 5.0 * (0.97 ** (k % n))


def make_W(N, seed):
    rng = np.random.default_rng(seed)
    U = rng.standard_normal((N, N))
    W = np.triu(U) + np.triu(U, 1).T
    return W.astype(np.float32)


def bytes_per_attempt(N, form):
    """Reference-algorithm bytes of one attempt: one J row, one q row (int8),
    8 bytes of random stream (site + uniform)."""
    if form == "fp8":
        return N * 1 + 4 + N * 1 + 8  # e4m3 row + its 4 B scale
    bJ = 4  # fp32 and bf16 hi+lo both 4 B per element
    return N * bJ + N * 1 + 8


def run_form(rocm, W, m, form, steps, warmup, seed, multi_flip=None, helper_wave=None):
    N = W.shape[0]
    prefs = dict(n_trotters=m, profile=1)
    if multi_flip is not None and form in ("fp32", "bf16x2"):
        prefs.update(multi_flip=multi_flip)
    if helper_wave is not None and form != "plain":
        prefs.update(helper_wave=helper_wave)
    if form == "plain":
        prefs.update(sweep="plain")
    else:
        prefs.update(sweep="local", field_precision=form)
    an = rocm.dense_graph_annealer(W, dtype=np.float32, **prefs)
    dev = rocm.device.get_active_device()
    an.seed(seed)
    an.prepare()
    an.randomize_spin()
    for k in range(warmup):
        an.anneal_one_step(schedule_G(k), BETA)
    an.randomize_spin()
    dev.synchronize()
    an.get_phase_times()  # reset the accumulators
    c0 = an.get_counters()
    t0 = time.perf_counter()
    for k in range(steps):
        an.anneal_one_step(schedule_G(k), BETA)
    dev.synchronize()
    t1 = time.perf_counter()
    phases = an.get_phase_times()
    c1 = an.get_counters()
    elapsed = t1 - t0
    attempts = steps * N * m
    sweep_ms, sweep_n = phases["sweep"]
    avg_sweep_ms = sweep_ms / max(sweep_n, 1)
    E = an.get_E()
    # the energy evaluation (calculate_E: integer-plane MFMA kernel, final
    # sums, copy of m doubles and a synchronisation), host wall time per call
    n_e = 10
    te = time.perf_counter()
    for _ in range(n_e):
        an.calculate_E()
    energy_call_ms = 1e3 * (time.perf_counter() - te) / n_e
    Eerr = an.get_energy_rounding_error()
    Eform = an.get_form_energy_error()
    flips = c1["flips"] - c0["flips"]
    batches = c1["batches"] - c0["batches"]
    res = {
        "attempts_per_s": attempts / elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "sweep_avg_ms": avg_sweep_ms,
        "acceptance": flips / attempts,
        "attempts_per_batch": attempts / max(batches, 1),
        "E_min": float(np.min(E)),
        "E_mean": float(np.mean(E)),
        "E_fp64_rounding_max_abs": Eerr,
        "E_form_max_abs_err": Eform,
        "E_form_max_rel_err": Eform / max(abs(float(np.min(E))), 1e-30),
        "flips_per_launch": flips / max(sweep_n, 1),
        "energy_call_ms": energy_call_ms,
    }
    if form != "plain":
        # the field GEMM's own time: HIP events around every phase, in a pass
        # after the timed region (the events would add gaps to the timed steps)
        an.set_preferences(profile=2)
        an.get_phase_times()
        for k in range(steps, steps + 10):
            an.anneal_one_step(schedule_G(k), BETA)
        dev.synchronize()
        g_ms, g_n = an.get_phase_times()["refresh"]
        if g_n:
            res["field_gemm_avg_ms"] = g_ms / g_n
    del an
    return res


def run_bipartite(rocm, steps, form, seed=1133557):
    """Bipartite SQA at N0 = N1 = 1024, m = 64, b0, b1, W ~ N(0,1) (fp32), the
    same schedule; m (N0 + N1) = 131,072 attempts per step."""
    N0 = N1 = 1024
    m = 64
    rng = np.random.default_rng(seed)
    b0 = rng.standard_normal(N0).astype(np.float32)
    b1 = rng.standard_normal(N1).astype(np.float32)
    W = rng.standard_normal((N1, N0)).astype(np.float32)
    an = rocm.bipartite_graph_annealer(b0, b1, W, dtype=np.float32, n_trotters=m, field_precision=form)
    dev = rocm.device.get_active_device()
    an.seed(1000)
    an.prepare()
    an.randomize_spin()
    for k in range(5):
        an.anneal_one_step(schedule_G(k), BETA)
    an.randomize_spin()
    dev.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        an.anneal_one_step(schedule_G(k), BETA)
    dev.synchronize()
    el = time.perf_counter() - t0
    E = an.get_E()
    return {"attempts_per_s": steps * m * (N0 + N1) / el, "ms_per_step": 1e3 * el / steps, "steps": steps,
            "E_min": float(np.min(E))}


def run_bf(rocm, N, seed=1133557):
    """Dense brute-force search over all 2^N x, W symmetric integer in [-8, 8]."""
    rng = np.random.default_rng(seed + N)
    U = rng.integers(-8, 9, size=(N, N)).astype(np.float32)
    W = np.triu(U) + np.triu(U, 1).T
    s = rocm.dense_graph_bf_searcher(W, dtype=np.float32)
    s.search()  # warm-up (allocations, code objects)
    t0 = time.perf_counter()
    s.search()
    el = time.perf_counter() - t0
    return {"N": N, "candidates_per_s": float(2 ** N) / el, "ms": 1e3 * el, "E_min": float(s.get_E()[0]),
            "n_solutions": len(s.get_x())}


def run_bg_bf(rocm, N0, N1, seed=1133557):
    """Bipartite brute-force search over all 2^(N0+N1) pairs, b0, b1, W
    integers in [-8, 8]."""
    rng = np.random.default_rng(seed + N0 * 64 + N1)
    b0 = rng.integers(-8, 9, N0).astype(np.float32)
    b1 = rng.integers(-8, 9, N1).astype(np.float32)
    W = rng.integers(-8, 9, (N1, N0)).astype(np.float32)
    s = rocm.bipartite_graph_bf_searcher(b0, b1, W, dtype=np.float32)
    s.search()  # warm-up
    t0 = time.perf_counter()
    s.search()
    el = time.perf_counter() - t0
    return {"N0": N0, "N1": N1, "pairs_per_s": float(2 ** (N0 + N1)) / el, "ms": 1e3 * el,
            "E_min": float(s.get_E()[0]), "n_solutions": len(s.get_x())}


def serve_jobs(n_jobs, seed):
    """70% dense / 30% bipartite (N0 = N1 = N/2) jobs, N log-uniform in
    [32, 8192], m uniform in {16..256}, weights N(0,1), each with its own
    schedule: 50..500 steps of G from G0 ~ U[2, 5] down to 0.01 geometrically,
    beta ~ U[20, 60]."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n_jobs):
        N = int(round(np.exp(rng.uniform(np.log(32), np.log(8192)))))
        m = int(rng.integers(16, 257))
        steps = int(rng.integers(50, 501))
        G0, beta = float(rng.uniform(2.0, 5.0)), float(rng.uniform(20.0, 60.0))
        G = G0 * (0.01 / G0) ** (np.arange(steps) / max(steps - 1, 1))
        if rng.random() < 0.7:
            U = rng.standard_normal((N, N)).astype(np.float32)
            W = np.triu(U) + np.triu(U, 1).T
            jobs.append(("dense", (W,), m, 5000 + i, G, beta, N * m * steps))
        else:
            h = max(N // 2, 1)
            jobs.append(("bipartite", (rng.standard_normal(h).astype(np.float32),
                                       rng.standard_normal(h).astype(np.float32),
                                       rng.standard_normal((h, h)).astype(np.float32)), m, 5000 + i, G, beta,
                         2 * h * m * steps))
    return jobs


def run_serve(rocm, n_jobs, load, seed=20240601):
    """Serving with continuous batching (sqaod_amd.rocm.Server): jobs are
    admitted in arrival order under a device-memory budget while others run,
    every tick steps all active jobs together (dense jobs in shared launches,
    bipartite jobs on parallel streams).  First all jobs are submitted at once
    to measure the server's capacity (jobs/s at saturation); then the same
    jobs arrive as a Poisson process at `load` times that rate, and latency
    = finish - arrival on the server's clock."""
    jobs = serve_jobs(n_jobs, seed)

    def submit(srv, j):
        kind, data, m, sd, G, beta, _ = j
        if kind == "dense":
            return srv.submit_dense(data[0], m, sd, G, beta)
        return srv.submit_bipartite(*data, m, sd, G, beta)

    srv = rocm.server()
    t0 = srv.now()
    for j in jobs:
        submit(srv, j)
    ticks = 0
    while srv.tick() > 0:
        ticks += 1
    makespan = srv.now() - t0
    done = srv.take_done()
    for d in done:
        srv.release(d)
    attempts = sum(j[-1] for j in jobs)
    capacity = n_jobs / makespan
    # Poisson arrivals at load x capacity
    rate = load * capacity
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    srv = rocm.server()
    start = srv.now()
    nxt, lat = 0, []

    def collect():
        for d in srv.take_done():
            r = srv.result(d)
            lat.append(r["finish_s"] - r["submit_s"])
            srv.release(d)

    while True:
        now = srv.now() - start
        while nxt < n_jobs and arrivals[nxt] <= now:
            submit(srv, jobs[nxt])
            nxt += 1
        st = srv.status()
        if st["pending"] + st["active"] == 0:
            if nxt >= n_jobs:
                break
            time.sleep(max(0.0, min(arrivals[nxt] - now, 0.05)))
            continue
        srv.tick()
        collect()
    collect()
    lat = np.array(lat) * 1e3
    return {"jobs": n_jobs, "dense_jobs": sum(j[0] == "dense" for j in jobs),
            "saturation_makespan_s": makespan, "saturation_ticks": ticks,
            "capacity_jobs_per_s": capacity, "attempts_per_s": attempts / makespan,
            "qps": rate, "load": load, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "note": "continuous batching: admission in arrival order under the memory budget, one tick steps "
                    "every active job; attempts_per_s at saturation (all jobs submitted at once)"}


def run_batch(rocm, n_inst, steps, seed=4242):
    """Batching: n_inst dense instances, N = round(64 * 32^u) (log-uniform in
    [64, 2048]), m uniform in {16, 32, 64, 128}, W symmetric i.i.d. U[-1, 1],
    bf16x2 fields, the schedule's first `steps` steps captured as one hipGraph
    after the autotuner has picked the launch shape."""
    rng = np.random.default_rng(seed)
    b = rocm.dense_batch(field_precision="bf16x2")
    shapes = []
    for i in range(n_inst):
        N = int(round(64 * 32 ** rng.random()))
        m = int(rng.choice([16, 32, 64, 128]))
        U = rng.uniform(-1.0, 1.0, (N, N)).astype(np.float32)
        W = np.triu(U) + np.triu(U, 1).T
        b.add_qubo(W, m, 1000 + i)
        shapes.append((N, m))
    t0 = time.perf_counter()
    b.prepare()
    t_prep = time.perf_counter() - t0
    G = np.array([schedule_G(k) for k in range(steps)])
    wpb, largest = b.autotune(G[0], BETA, 2)
    t0 = time.perf_counter()
    ms = b.anneal(G, BETA, graph=True)
    wall = time.perf_counter() - t0
    att = b.attempts_per_step() * steps
    return {"instances": n_inst, "steps": steps, "attempts_per_step": b.attempts_per_step(),
            "attempts_per_s": att / (ms * 1e-3), "ms_per_step": ms / steps, "wall_s_incl_capture": wall,
            "prepare_s": t_prep, "waves_per_block": wpb, "order": "largest_first" if largest else "given",
            "sum_N_m": int(sum(N * m for N, m in shapes)), "max_N": max(N for N, _ in shapes)}


def cpu_baseline(W, m, seconds, seed):
    """sqaod.cpu OpenMP coloring annealer (C restatement) on the host cores."""
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0))
    env_t = os.environ.get("OMP_NUM_THREADS")
    if env_t and env_t.isdigit():
        threads = min(threads, int(env_t))
    N = W.shape[0]
    ora = O.DenseGraphAnnealer(np.float32, threads)
    ora.seed(seed)
    ora.set_qubo(W)
    ora.set_m(m)
    ora.prepare()
    ora.randomize_spin()
    ora.anneal_one_step(schedule_G(0), BETA)  # warm-up
    steps = 0
    t0 = time.perf_counter()
    while True:
        ora.anneal_one_step(schedule_G(steps + 1), BETA)
        steps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    return {
        "value": steps * N * m / el,
        "unit": "spin-flip attempts/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} annealOneStep(s) of the OpenMP coloring annealer at N={N} m={m} "
                  f"(steps 1..{steps} of the schedule after 1 warm-up step), {threads} threads, "
                  f"gcc -O2 -msse2 SSE2-order dot",
    }


def copy_calibration():
    """A device copy of known size (1 GiB) for scaling rocprof FETCH_SIZE/WRITE_SIZE."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n))
    hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n))
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=205, help="timed annealOneStep calls per form")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--N", type=int, default=4096)
    ap.add_argument("--m", type=int, default=256)
    ap.add_argument("--forms", default="bf16x2,fp32,fp8,plain")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--multi-flip", default=None,
                    help="multi-flip batches of the float local-field forms: auto, on, off or the "
                         "threshold 2..64 (default: the library's, auto)")
    ap.add_argument("--helper-wave", default=None, choices=["on", "off"],
                    help="second wave per slice applying the deferred rows (default: the library's, on)")
    ap.add_argument("--no-extras", action="store_true", help="skip the bipartite, BF search and serving runs")
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--batch-instances", type=int, default=512)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the dense SQA path runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import sqaod_amd.rocm as rocm

    if args.copy_calibration:
        copy_calibration()

    N, m = args.N, args.m
    W = make_W(N, 1133557)
    forms = [f for f in args.forms.split(",") if f]
    results = {}
    for form in forms:
        results[form] = run_form(rocm, W, m, form, args.steps, args.warmup, seed=1000, multi_flip=args.multi_flip,
                                 helper_wave=args.helper_wave)
        print(f"# {form}: {results[form]['attempts_per_s']:.4g} attempts/s, "
              f"{results[form]['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)

    head = "bf16x2" if "bf16x2" in results else forms[0]
    r = results[head]
    B_att = bytes_per_attempt(N, head)
    ref_per_launch = B_att * N * m
    sweep_s = r["sweep_avg_ms"] * 1e-3
    algo_per_launch = sweep_bytes_per_launch(N, m, head, r["flips_per_launch"]) if head != "plain" else ref_per_launch
    achieved = algo_per_launch / sweep_s / 1e9
    traffic = measured_traffic(N, m, head, args.steps, args.warmup)
    window = (f"schedule steps 0-{args.steps - 1} of 205 (after {args.warmup} warm-up steps and re-randomized "
              f"spins)")
    out = {
        "metric": "spin-flips/s dense SQA N=4096 m=256 bf16-split vs fp32; serve p50/p99 ms",
        "value": r["attempts_per_s"],
        "unit": "spin-flip attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: W symmetric i.i.d. N(0,1) from a seeded generator",
        "config": {
            "workload": f"dense SQA annealOneStep, N={N}, m={m}, G=5.0*0.97^k, beta=50, "
                        f"local fields, J as bf16 hi+lo, Philox streams; timed: {window}",
            "N": N, "m": m, "form": head, "window": window,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_frac": (traffic / sweep_s / 1e9 / HBM_PEAK_GBS) if traffic else None,
            "kernel": "denseSweepKernel",
            "algorithmic_bytes_per_launch": algo_per_launch,
            "note": "achieved = the local-field sweep's algorithmic bytes per launch (stream, decisions, fields, "
                    "spins, one J row per accepted flip at the measured flip count) / the sweep's average "
                    "duration (HIP events on the solver's stream); traffic = measured HBM bytes per launch of "
                    "this window (rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/), traffic_frac = traffic / "
                    "duration / peak",
        },
        "reference_work_equivalent": {
            "bytes_per_launch": ref_per_launch,
            "GBps": ref_per_launch / sweep_s / 1e9,
            "note": "N*m*(4N+N+8): the bytes the reference algorithm (a J-row . q-row dot per attempt) "
                    "would move for the same attempts; not bytes this sweep moves",
        },
        "forms": results,
    }
    gemm_ms = r.get("field_gemm_avg_ms")
    if gemm_ms:
        planes = {"bf16x2": 2, "fp32": 3, "fp8": 1}[head]
        flops = 2.0 * m * N * N
        peak = FP8_MFMA_PEAK_TFLOPS if head == "fp8" else BF16_MFMA_PEAK_TFLOPS
        out["field_gemm"] = {
            "kernel": "fieldGemm (F = Q J on MFMA)", "avg_ms": gemm_ms,
            "algorithmic_TFLOPs": flops / (gemm_ms * 1e-3) / 1e12,
            "mfma_issued_TFLOPs": planes * flops / (gemm_ms * 1e-3) / 1e12,
            "peak_TFLOPs": peak,
            "mfma_frac": planes * flops / (gemm_ms * 1e-3) / 1e12 / peak,
            "note": f"2 m N^2 = {flops:.4g} flop per refresh, issued {planes}x on "
                    f"{'fp8' if head == 'fp8' else 'bf16'} MFMA ({planes} J planes); time from HIP events around "
                    f"the refresh in 10 steps after the timed region",
        }
    out["energy"] = {
        "kernel": "planeEnergyKernel + planeFinalKernel (E of all m slices, J as int8 digit planes on MFMA)",
        "call_ms": r["energy_call_ms"],
        "note": "host wall time per calculate_E, including its device-to-host copy of m doubles and the "
                "stream synchronisation; the kernels' own durations are in the rocprofv3 summary",
    }
    if "plain" in results and head in results:
        out["speedup_vs_plain_fp32"] = r["attempts_per_s"] / results["plain"]["attempts_per_s"]
    if "fp32" in results and head in results:
        out["speedup_vs_local_fp32"] = r["attempts_per_s"] / results["fp32"]["attempts_per_s"]
    if not args.no_extras:
        out["bipartite"] = {f: run_bipartite(rocm, args.steps, f) for f in ("fp32", "bf16x2")}
        print(f"# bipartite: {out['bipartite']['fp32']['attempts_per_s']:.4g} attempts/s", file=sys.stderr,
              flush=True)
        out["batch"] = run_batch(rocm, args.batch_instances, 200)
        print(f"# batch: {out['batch']['attempts_per_s']:.4g} attempts/s over {args.batch_instances} instances",
              file=sys.stderr, flush=True)
        out["bf_search"] = [run_bf(rocm, n) for n in (24, 30)]
        out["bipartite_bf_search"] = [run_bg_bf(rocm, 12, 12), run_bg_bf(rocm, 16, 16)]
        print(f"# bipartite bf search 16x16: {out['bipartite_bf_search'][1]['pairs_per_s']:.4g} pairs/s",
              file=sys.stderr, flush=True)
        print(f"# bf search N=30: {out['bf_search'][1]['candidates_per_s']:.4g} candidates/s", file=sys.stderr,
              flush=True)
        out["serve"] = run_serve(rocm, args.serve_jobs, 0.7)
        print(f"# serve: p50 {out['serve']['p50_ms']:.1f} ms, p99 {out['serve']['p99_ms']:.1f} ms", file=sys.stderr,
              flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(W, m, args.cpu_seconds, seed=1000)
    print(json.dumps(out), flush=True)
    # release every device object while the runtime (and any profiler attached
    # to it) is fully up, rather than during interpreter teardown
    import gc
    gc.collect()
    rocm.device.unload()
    gc.collect()


if __name__ == "__main__":
    main()
