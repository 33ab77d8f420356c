"""Benchmark of the sparsediag hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): spin-1/2
Heisenberg ring L=24, J=1, Sz=0 sector, D = C(24,12) = 2,704,156 states,
Lanczos vectors stored in FP32 (FP64 accumulation).  A step is one
matrix-free H.v over the whole sector, y = H x; `value` is the H.v
throughput in GB/s counted with the algorithmic bytes 8*D per product
(read x once, write y once, FP32), with x already resident in HBM.

Also reported (legs, each selectable with --only): the device-resident
Lanczos step; the L=24 eigenvalues with FP32 and FP64 storage against the
oracle's FP64 run (tests/golden/ring24_oracle.json); the 5x6 J1-J2 lattice
(H.v and 300 fixed iterations); the H.v of an L=24 momentum sector; 1024
random XXZ chains as one batch; a Poisson serving stream at a sweep of
offered rates with every result checked against an FP64 re-solve; the
sweep of every (Sz, k) sector of a ring; and a CPU baseline (the oracle's
reference-style restatement on the host cores: H.v on one and on all
threads, a phase-timed L=16 solve, sequential and farmed batch solves, and
a single-worker FIFO on the serving trace).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of hv_kernel<float,false> on this workload (L=24), per
# schedule (the order of SCHEDULES below), from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE scaled by a copy kernel of known size
# (profiles/hv_kernel_heisenberg_L24.md); valid for that kernel version and L
# only.  The schedules that spill write their scratch too.
PMC_TRAFFIC_BYTES = {24: [53.38e6, 53.47e6, 53.35e6, 53.30e6, 58.25e6, 54.01e6]}
# SHA-256 of the row kernel's source (alps_amd/csrc/ed_device.hpp) when those
# counters were taken: if the source differs, the table no longer describes
# the kernel that runs, and the bench reports traffic as unknown instead.
PMC_TRAFFIC_KERNEL_SHA256 = "8dd16111070282a21ca2d360f4733c5027e3ac0292ebee8e726dea2ee30cccc5"
PMC_TRAFFIC_SOURCE = ("profiles/hv_kernel_heisenberg_L24.md, section 'HBM byte counters', column "
                      "'traffic per launch' (FETCH_SIZE + WRITE_SIZE, scaled by the copy kernel), "
                      "row of the chosen schedule")
# H.v schedules (waves per SIMD, low-low / high-high / cross chunk sizes), in
# the order of the engine's schedule table
SCHEDULES = [(8, 4, 4, 2), (8, 4, 8, 2), (6, 4, 8, 2), (6, 4, 4, 2), (8, 8, 4, 2), (6, 8, 8, 4)]


class HipEvents:
    """hipEvent timing on the null stream, which is the stream the engine
    launches on by default."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.c_void_p]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e):
        assert self.hip.hipEventRecord(e, None) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0


def heisenberg_ring(L):
    import numpy as np

    import alps_amd as A

    s = np.arange(L)
    return A.SpinModel(L, s, (s + 1) % L, 1.0, 1.0, L // 2)


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return model, os.cpu_count(), usable


def cpu_baseline(L, rows, budget_s=20.0):
    """The oracle (the reference's algorithm restated: map-of-maps matrix by
    binary-search lookup, ietl Lanczos with the T-matrix test every step) on
    the host cores, on bounded samples of each workload."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    model, ncpu, usable = cpu_info()
    threads = max(1, min(usable, 16))
    om = O.lattice_model("chain lattice", L)
    D = math.comb(L, L // 2)
    hv = {}
    for th in (1, threads):
        r = O.bench_hv(om, L // 2, rows, reps=3, threads=th)
        t_full = r["seconds"] * D / rows  # rows cost the same on average
        hv[th] = (8.0 * D / t_full / 1e9, r)
    gbs1, r1 = hv[1]
    # C1: the L=16 ring, Sz=0, at most 200 nlowest iterations, two levels
    c1 = O.sparsediag_timed(O.lattice_model("chain lattice", 16), 8, max_iter=200, nev=2, seed=1)
    # C4: the batch's problems one at a time (the reference's task mode), and
    # farmed over the host threads (ctypes releases the GIL)
    models = xxz_batch(1024, 1, list(range(12, 21)))

    def oracle_model(m):
        return O.Model(m.n_sites, m.bond_site1, m.bond_site2, m.jz, m.jxy, h=m.h)

    seq, t0 = 0, time.perf_counter()
    while seq < len(models) and time.perf_counter() - t0 < budget_s / 4:
        O.sparsediag(oracle_model(models[seq]), models[seq].n_up, nev=1, seed=seq)
        seq += 1
    t_seq = time.perf_counter() - t0
    farm_n = max(seq, 1) * threads
    t1 = time.perf_counter()
    with ThreadPoolExecutor(threads) as pool:
        list(pool.map(lambda i: O.sparsediag(oracle_model(models[i]), models[i].n_up, nev=1, seed=i),
                      range(min(farm_n, len(models)))))
    t_farm = time.perf_counter() - t1
    # C5: one worker serving the stream's requests first come first served,
    # on the prefix of the arrival trace that fits the budget; requests of
    # more than 2e5 states are not run on the CPU (counted as skipped)
    smodels, arrivals = serve_stream(400, 2, 200.0)
    svc, arr, skipped, t2 = [], [], 0, time.perf_counter()
    for m, a in zip(smodels, arrivals):
        if time.perf_counter() - t2 > budget_s / 2:
            break
        if math.comb(m.n_sites, m.n_up) > 200000:
            skipped += 1
            continue
        s0 = time.perf_counter()
        O.sparsediag(oracle_model(m), m.n_up, nev=1, seed=0)
        svc.append(time.perf_counter() - s0)
        arr.append(a)
    free_at, lat = 0.0, []
    for a, sv in zip(arr, svc):
        start = max(a, free_at)
        free_at = start + sv
        lat.append(free_at - a)
    lat = np.array(lat) if lat else np.zeros(1)
    return {
        "value": round(gbs1, 4),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "cpu_model": model, "host_cpus": ncpu, "usable_cpus": usable,
        "sample": (f"oracle map-of-maps H.v (reference storage and row order) over rows [0,{rows}) of the "
                   f"L={L} Sz=0 matrix ({r1['nnz']} stored elements, built in {r1['build_seconds']:.1f} s by "
                   f"binary-search lookup), median of 3, scaled to D={D}; GB/s counted as 8*D bytes per H.v"),
        "hv_gbs_1_thread": round(gbs1, 4),
        "hv_gbs_all_threads": round(hv[threads][0], 4), "threads": threads,
        "c1_solve": {"workload": "Heisenberg ring L=16, Sz=0, nlowest Lanczos, at most 200 iterations, 2 levels, "
                                 "1 thread", "dim": c1["dim"], "iterations": c1["iterations"],
                     "eigenvalues": c1["ev"].tolist(), "gap": float(c1["ev"][1] - c1["ev"][0]),
                     "phase_s": {k: round(v, 5) for k, v in c1["phases"].items()},
                     "total_s": round(sum(c1["phases"].values()), 4)},
        "c4_sequential_solves_per_s": round(seq / t_seq, 2), "c4_sequential_sample": seq,
        "c4_farmed_solves_per_s": round(min(farm_n, len(models)) / t_farm, 2),
        "c4_farmed_threads": threads,
        "c5_fifo": {"requests_run": len(svc), "skipped_over_2e5_states": skipped,
                    "mean_service_ms": round(1e3 * float(np.mean(svc)) if svc else 0.0, 2),
                    "max_solves_per_s": round(1.0 / float(np.mean(svc)), 2) if svc else None,
                    "p50_latency_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                    "p99_latency_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
                    "workload": "the serving stream's arrival trace at 200/s (prefix that fits the budget), one "
                                "worker, first come first served, 1 thread"},
    }


def xxz_batch(n, seed, Ls, field=False, any_sector=False):
    """Random-bond XXZ chains: L uniform over Ls, J_i ~ U[0.5,1.5],
    Delta ~ U[0,2] (Jz_i = Delta J_i, Jxy_i = J_i), open or periodic with
    probability 1/2, optional uniform field h ~ U[0,0.5].  Sector Sz=0, or
    Sz=+1/2 for odd L; with any_sector, n_up uniform over 0..L."""
    import numpy as np

    import alps_amd as A

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        L = int(rng.choice(Ls))
        periodic = bool(rng.integers(2))
        nb = L if periodic else L - 1
        s = np.arange(nb)
        J = rng.uniform(0.5, 1.5, nb)
        delta = rng.uniform(0.0, 2.0)
        h = float(rng.uniform(0.0, 0.5)) if field else 0.0
        n_up = int(rng.integers(0, L + 1)) if any_sector else (L + 1) // 2
        out.append(A.SpinModel(L, s, (s + 1) % L, delta * J, J, n_up, h=h))
    return out


def batch_leg(n, seq_sample, seed=1, graphs=False):
    """1024 random XXZ chains (L in 12..20) solved in one batch, against the
    same solves one at a time (alps_ed_sparsediag) on a sample."""
    import numpy as np

    import alps_amd as A

    models = xxz_batch(n, seed, list(range(12, 21)))
    free = A.device_info()["hbm_bytes"]
    b = A.Batch("f32", hbm_bytes=int(free * 0.75), steps_per_round=8, host_threads=12, graphs=graphs)
    t0 = time.perf_counter()
    tickets = [b.submit(m, nev=1, seed=i) for i, m in enumerate(models)]
    b.run()
    t_batch = time.perf_counter() - t0
    st = b.stats()
    res = [b.result(t) for t in tickets]
    b.close()
    # the FP64 batch on the same problems is the accuracy reference
    b64 = A.Batch("f64", hbm_bytes=int(free * 0.5), steps_per_round=8, host_threads=12)
    t1 = time.perf_counter()
    t64 = [b64.submit(m, nev=1, seed=i) for i, m in enumerate(models)]
    b64.run()
    t_batch64 = time.perf_counter() - t1
    r64 = [b64.result(t) for t in t64]
    b64.close()
    rel = max(abs(r["ev"][0] - q["ev"][0]) / abs(q["ev"][0]) for r, q in zip(res, r64))
    # one solve at a time on a sample of the same problems
    t2 = time.perf_counter()
    for i in range(seq_sample):
        A.sparsediag(models[i], nev=1, seed=i, dtype="f32")
    t_seq = (time.perf_counter() - t2) / seq_sample
    dims = [r["dim"] for r in res]
    return {
        "solves": n,
        "states_total": int(sum(dims)),
        "batch_s": round(t_batch, 4),
        "solves_per_s": round(n / t_batch, 1),
        "fp64_batch_solves_per_s": round(n / t_batch64, 1),
        "one_at_a_time_solves_per_s": round(1.0 / t_seq, 2),
        "speedup_vs_one_at_a_time": round((n / t_batch) * t_seq, 2),
        "one_at_a_time_sample": seq_sample,
        "max_rel_err_f32_vs_f64": rel,
        "mean_iterations": round(float(np.mean([r["iterations"] for r in res])), 1),
        "rounds": st["rounds"],
        "seconds_split": {k: round(st["seconds_" + k], 4) for k in ("admit", "steps", "decide", "refine")},
        "hv_rows_f32": st["hv_rows_f32"],
        "graphs": bool(graphs), "graph_rounds": st["graph_rounds"],
        "workload": "random-bond XXZ chains, L uniform 12..20, J_i~U[0.5,1.5], Delta~U[0,2], open/periodic 1:1, "
                    "Sz=0 (Sz=1/2 for odd L), lowest eigenvalue, FP32 vectors + FP64 refinement",
    }


def serve_stream(n, seed, rate):
    """Poisson request stream: XXZ chains, L uniform 12..26, field h ~
    U[0,0.5], sector uniform over n_up = 0..L; arrival times at `rate`/s."""
    import numpy as np

    models = xxz_batch(n, seed, list(range(12, 27)), field=True, any_sector=True)
    rng = np.random.default_rng(seed + 1)
    return models, np.cumsum(rng.exponential(1.0 / rate, n))


def serve_run(models, arrivals, graphs=False):
    """Continuous batching of one arrival trace: requests join the batch as
    they arrive; returns per-request latency (arrival to result), the
    results and the wall time."""
    import numpy as np

    import alps_amd as A

    n = len(models)
    free = A.device_info()["hbm_bytes"]
    b = A.Batch("f32", hbm_bytes=int(free * 0.75), steps_per_round=4, host_threads=12, graphs=graphs)
    t0 = time.perf_counter()
    i = 0
    arrival_of, lat, res = {}, {}, {}
    last_print = t0
    while len(lat) < n:
        now = time.perf_counter() - t0
        while i < n and arrivals[i] <= now:
            arrival_of[b.submit(models[i], nev=1, seed=i)] = (i, arrivals[i])
            i += 1
        a, p = b.progress()
        if a == 0 and p == 0 and i < n:
            time.sleep(max(0.0, arrivals[i] - (time.perf_counter() - t0)))
        done_t = time.perf_counter() - t0
        for t in b.completed():
            k, arr = arrival_of[t]
            lat[k] = done_t - arr
            res[k] = b.result(t)
        if time.perf_counter() - last_print > 20:
            print(f"serve: {len(lat)}/{n} done", file=sys.stderr, flush=True)
            last_print = time.perf_counter()
    wall = time.perf_counter() - t0
    st = b.stats()
    b.close()
    return np.array([lat[k] for k in range(n)]), [res[k] for k in range(n)], wall, st


def serve_leg(n, rates, seed=2, graphs=False, seq_sample=48):
    """A Poisson request stream at a sweep of offered rates up to saturation
    (continuous batching with HBM admission), every served result checked
    against an FP64 re-solve of the same request, and the same requests one
    at a time through alps_ed_sparsediag."""
    import numpy as np

    import alps_amd as A

    sweep = []
    worst = 0.0
    for rate in rates:
        models, arrivals = serve_stream(n, seed, rate)
        lat, res, wall, st = serve_run(models, arrivals, graphs)
        # FP64 re-solve of every request (one batch, not timed)
        free = A.device_info()["hbm_bytes"]
        b64 = A.Batch("f64", hbm_bytes=int(free * 0.5), steps_per_round=8, host_threads=12)
        t64 = [b64.submit(m, nev=1, seed=i) for i, m in enumerate(models)]
        b64.run()
        r64 = [b64.result(t) for t in t64]
        b64.close()
        rel = []
        for r, q in zip(res, r64):
            if len(q["ev"]) == 0:
                rel.append(0.0 if len(r["ev"]) == 0 else float("inf"))
            else:
                rel.append(abs(r["ev"][0] - q["ev"][0]) / max(abs(q["ev"][0]), 1e-300))
        worst = max(worst, max(rel))
        sweep.append({"offered_rate_per_s": rate, "solves_per_s": round(n / wall, 1),
                      "p50_latency_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                      "p99_latency_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
                      "wall_s": round(wall, 3), "rounds": st["rounds"], "graph_rounds": st["graph_rounds"],
                      "max_rel_err_vs_fp64": max(rel), "results_over_1e-10": int(sum(e > 1e-10 for e in rel))})
    # one at a time on a sample of the same requests
    models, _ = serve_stream(n, seed, rates[0])
    t0 = time.perf_counter()
    for i in range(seq_sample):
        A.sparsediag(models[i], nev=1, seed=i, dtype="f32")
    one = seq_sample / (time.perf_counter() - t0)
    best = max(e["solves_per_s"] for e in sweep)
    first = sweep[0]
    return {
        "requests_per_rate": n, "graphs": bool(graphs),
        "offered_rate_per_s": first["offered_rate_per_s"], "solves_per_s": first["solves_per_s"],
        "p50_latency_ms": first["p50_latency_ms"], "p99_latency_ms": first["p99_latency_ms"],
        "saturation_solves_per_s": best,
        "one_at_a_time_solves_per_s": round(one, 2), "one_at_a_time_sample": seq_sample,
        "speedup_at_saturation_vs_one_at_a_time": round(best / one, 1),
        "max_rel_err_vs_fp64_all_rates": worst,
        "rate_sweep": sweep,
        "workload": "Poisson arrivals; XXZ chains L uniform 12..26, Delta~U[0,2], h~U[0,0.5], n_up uniform 0..L, "
                    "FP32 vectors + FP64 refinement, HBM-footprint admission; every result against an FP64 "
                    "re-solve",
    }


def j1j2_leg(ev, iterations=300):
    """5x6 J1-J2 square lattice (periodic, J1=1, J2=0.5, Sz=0,
    D=155,117,520): H.v rate and a fixed-length Lanczos run with FP32
    vectors (basis kept in HBM, FP64 refinement) next to FP64 vectors."""
    import numpy as np

    import alps_amd as A

    (t,) = A.describe_parameter_file('LATTICE="frustrated square lattice"; L=6; W=5; J=1; J1=0.5; '
                                     "Sz_total=0; {}")
    m = A.SpinModel(t["n_sites"], np.array(t["bond_s"]), np.array(t["bond_t"]), np.array(t["jz"]),
                    np.array(t["jxy"]), 15)
    op = A.Operator(m)
    sched, sched_s = op.schedule()
    D = op.dim
    x, y = A.Vector(D, "f32"), A.Vector(D, "f32")
    x.fill_random(1)
    for _ in range(2):
        op.matvec(x, y)
    ev.sync()
    e0, e1 = ev.event(), ev.event()
    ev.record(e0)
    for _ in range(10):
        op.matvec(x, y)
    ev.record(e1)
    ms = ev.elapsed_ms(e0, e1) / 10
    del x, y
    r32 = A.sparsediag(m, max_iter=iterations, fixed=True, nev=2, seed=1, dtype="f32")
    r64 = A.sparsediag(m, max_iter=iterations, fixed=True, nev=2, seed=1, dtype="f64")
    rel = (np.abs(r32["ev"] - r64["ev"]) / np.abs(r64["ev"])).tolist()
    return {
        "workload": "J1-J2 Heisenberg, 6x5 periodic square lattice, J1=1, J2=0.5, Sz=0",
        "dim": D, "bonds": len(t["bond_s"]),
        "hv_ms": round(ms, 4), "hv_gbs_8D": round(8.0 * D / (ms * 1e-3) / 1e9, 1),
        "hv_frac_of_hbm_peak": round(8.0 * D / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "hv_schedule": {"chosen": sched, "waves_cll_chh_cx": SCHEDULES[sched],
                        "ms_per_schedule": None if sched_s is None else [round(t * 1e3, 3) for t in sched_s]},
        "iterations": iterations,
        "f32_eigenvalues": r32["ev"].tolist(), "f64_eigenvalues": r64["ev"].tolist(), "rel_err_f32": rel,
        "solve_s_f32": round(r32["seconds"], 3), "solve_s_f64": round(r64["seconds"], 3),
    }


def bloch_leg(ev, L=24, reps=50):
    """H.v in the momentum sector k = 0 of the Heisenberg ring's Sz=0 sector
    (the reference's default mode on periodic lattices): realified FP32
    vectors of 2 d elements."""
    import numpy as np

    import alps_amd as A

    s = np.arange(L)
    m = A.SpinModel(L, s, (s + 1) % L, 1.0, 1.0, L // 2)
    perms = np.array([[(i + j) % L for i in range(L)] for j in range(L)])
    op = A.Operator(m, translations=(np.ones(L, dtype=complex), perms))
    n = op.dim
    x, y = A.Vector(n, "f32"), A.Vector(n, "f32")
    x.fill_random(1)
    for _ in range(5):
        op.matvec(x, y)
    ev.sync()
    e0, e1 = ev.event(), ev.event()
    ev.record(e0)
    for _ in range(reps):
        op.matvec(x, y)
    ev.record(e1)
    ms = ev.elapsed_ms(e0, e1) / reps
    return {"workload": f"Heisenberg ring L={L}, Sz=0, momentum k=0 (Bloch basis), FP32 realified vectors",
            "representatives": n // 2, "vector_length": n, "hv_us": round(ms * 1e3, 3),
            "hv_gbs_8n": round(8.0 * n / (ms * 1e-3) / 1e9, 1),
            "hv_frac_of_hbm_peak": round(8.0 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def sweep_leg(L=20, sample=42):
    """The parameter-file task of a periodic XXZ ring without a fixed Sz:
    every (Sz, momentum) sector (DiagMatrix::dostep's loops), run as one
    batch, against a sample of the same sectors solved one at a time through
    the single-sector path, scaled to all of them."""
    import numpy as np

    import alps_amd as A

    text = f'LATTICE="chain lattice"; L={L}; J=1; Jz=0.7; NUMBER_EIGENVALUES=2; {{}}'
    A.run_parameter_file(text)  # first sweep of the process: caches and arena
    t0 = time.perf_counter()
    (t,) = A.run_parameter_file(text)
    batched = time.perf_counter() - t0
    (d,) = A.describe_parameter_file(text)
    ks = d["momenta"]
    n_sectors = (L + 1) * len(ks)
    s1, s2 = np.array(d["bond_s"]), np.array(d["bond_t"])
    jz, jxy = np.array(d["jz"]), np.array(d["jxy"])
    # translations of the ring: shift by j sites, phase exp(i k j)
    perms = np.array([[(i + j) % L for i in range(L)] for j in range(L)])
    rng = np.random.default_rng(5)
    picks = rng.choice(n_sectors, size=min(sample, n_sectors), replace=False)
    t1 = time.perf_counter()
    for p in picks:
        n_up, q = divmod(int(p), len(ks))
        k = ks[q][0]
        phases = np.exp(1j * k * np.arange(L))
        A.sparsediag(A.SpinModel(L, s1, s2, jz, jxy, n_up), nev=2, translations=(phases, perms))
    one = (time.perf_counter() - t1) / len(picks) * n_sectors
    return {"workload": f"XXZ ring L={L}, Jz=0.7, every Sz and momentum sector, lowest two levels, FP64",
            "sectors": n_sectors, "ground_state_energy": t["ground_state_energy"],
            "batched_s": round(batched, 4), "one_at_a_time_s": round(one, 3), "one_at_a_time_sample": len(picks),
            "speedup": round(one / batched, 1)}


def traffic_table_valid(L):
    """Whether PMC_TRAFFIC_BYTES describes the kernel that runs: a row for L
    and the row kernel's source unchanged since the counters were taken."""
    import hashlib

    if L not in PMC_TRAFFIC_BYTES:
        return False, f"no byte counters recorded for L={L}"
    with open(os.path.join(ROOT, "alps_amd", "csrc", "ed_device.hpp"), "rb") as f:
        sha = hashlib.sha256(f.read()).hexdigest()
    if sha != PMC_TRAFFIC_KERNEL_SHA256:
        return False, "row kernel source changed since the byte counters were taken; rerun them"
    return True, None


def eigen_leg(model, L, ev, op):
    """The L-ring's lowest levels with FP32 and FP64 vector storage, against
    the oracle's FP64 sparsediag of the same sector (stored for L=24 in
    tests/golden/ring24_oracle.json); the FP32 single solve's wall time
    against its device step time (iterations x the Lanczos step)."""
    import numpy as np

    import alps_amd as A

    ref = {}
    path = os.path.join(ROOT, "tests", "golden", "ring24_oracle.json")
    if L == 24 and os.path.exists(path):
        with open(path) as f:
            for r in json.load(f)["runs"]:
                if r["n_up"] == 12:
                    ref[r["nev"]] = r["eigenvalues"]
    # Lanczos step time of the FP32 recurrence without reorthogonalization
    lz = A.Lanczos(op, "f32", max_steps=130)
    lz.start(seed=1)
    ev.sync()
    e0, e1 = ev.event(), ev.event()
    ev.record(e0)
    lz.steps(100)
    ev.record(e1)
    step_ms = ev.elapsed_ms(e0, e1) / 100
    lz.close()
    out = {"lanczos_step_ms_f32": round(step_ms, 5), "runs": []}
    for nev in (1, 2, 4):
        for dtype in ("f32", "f64"):
            A.sparsediag(model, nev=nev, seed=1, dtype=dtype)  # first call of a shape: allocations, tuning
            r = A.sparsediag(model, nev=nev, seed=1, dtype=dtype)
            e = {"nev": nev, "dtype": dtype, "storage": r["storage"], "iterations": r["iterations"],
                 "reorthogonalized_steps": r["reorthogonalizations"], "solve_ms": round(r["seconds"] * 1e3, 3),
                 "eigenvalues": r["ev"].tolist()}
            if nev in ref and len(ref[nev]) == len(r["ev"]):
                e["max_rel_err_vs_oracle"] = float(np.max(np.abs(r["ev"] - ref[nev]) / np.abs(ref[nev])))
            if dtype == "f32" and nev == 1:
                e["solve_over_device_steps"] = round(r["seconds"] * 1e3 / (r["iterations"] * step_ms), 3)
            out["runs"].append(e)
    return out


LEGS = ("hv", "eigen", "j1j2", "bloch", "batch", "serve", "sweep", "cpu")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--L", type=int, default=24)
    ap.add_argument("--only", default="", help="comma-separated legs to run (of %s); default all" % ",".join(LEGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-eigen", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=300000)
    ap.add_argument("--no-j1j2", action="store_true", help="skip the 5x6 J1-J2 leg (about 200 GB of HBM)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-size", type=int, default=1024)
    ap.add_argument("--seq-sample", type=int, default=64)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rates", default="200,1000,3000,10000,40000")
    ap.add_argument("--graphs", action="store_true", help="batch and serve legs replay repeated rounds from HIP graphs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported (only --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    only = set(filter(None, args.only.split(",")))
    unknown = only - set(LEGS)
    if unknown:
        print(f"bench.py: unknown legs {sorted(unknown)} (known: {', '.join(LEGS)})", file=sys.stderr)
        sys.exit(2)
    skip = {"eigen": args.no_eigen, "j1j2": args.no_j1j2, "batch": args.no_batch, "serve": args.no_serve,
            "sweep": args.no_sweep, "cpu": args.no_cpu_baseline}

    def want(leg):
        return leg in only if only else not skip.get(leg, False)

    import numpy as np

    import alps_amd as A

    ev = HipEvents()
    info = A.device_info()
    L = args.L
    model = heisenberg_ring(L)
    op = A.Operator(model)  # times the FP32 H.v schedules on this sector, keeps the fastest
    sched, sched_s = op.schedule()
    D = op.dim
    x = A.Vector(D, "f32")
    y = A.Vector(D, "f32")
    x.fill_random(1)
    bytes_per_hv = 8.0 * D

    for _ in range(args.warmup):
        op.matvec(x, y)
    ev.sync()
    e0, e1 = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(e0)
    for _ in range(args.steps):
        op.matvec(x, y)
    ev.record(e1)
    ms = ev.elapsed_ms(e0, e1)
    ev.sync()
    wall = time.perf_counter() - t0
    ms_per_hv = ms / args.steps
    gbs = bytes_per_hv / (ms_per_hv * 1e-3) / 1e9
    traffic_ok, traffic_note = traffic_table_valid(L)

    out = {
        "metric": "H·v GB/s (% HBM peak); batched solves/sec + p99 latency at 1e-10 rel err",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_hv, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 vectors / f64 accumulate",
        "data": "synthetic",
        "config": {"workload": f"Heisenberg ring L={L}, J=1, Sz=0, D={D}, matrix-free H.v, FP32 vectors",
                   "L": L, "dim": D, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": PMC_TRAFFIC_BYTES[L][sched] if traffic_ok else None,
                     "traffic_source": PMC_TRAFFIC_SOURCE if traffic_ok else traffic_note,
                     "kernel": "hv_kernel<float,false,%d,%d,%d,%d>" % SCHEDULES[sched],
                     "bytes_per_launch": bytes_per_hv},
        "timed_wall_s": round(wall, 4),
        "hv_schedule": {"chosen": sched, "waves_cll_chh_cx": SCHEDULES[sched],
                        "us_per_schedule": None if sched_s is None else [round(t * 1e6, 2) for t in sched_s]},
        "device": info["name"],
    }
    if want("hv"):
        # plain device copy of the same FP32 vector (4*D read + 4*D written):
        # the practical streaming rate, and the reference kernel of known
        # size for scaling the HBM byte counters of a profile
        nc = max(1, min(args.steps, 50))
        for _ in range(3):
            y.copy_from(x)
        ev.sync()
        e4, e5 = ev.event(), ev.event()
        ev.record(e4)
        for _ in range(nc):
            y.copy_from(x)
        ev.record(e5)
        out["copy_gbs"] = round(bytes_per_hv / (ev.elapsed_ms(e4, e5) / nc * 1e-3) / 1e9, 2)
        # device-resident Lanczos step (H.v with alpha + update with beta)
        nl = min(args.steps, 400)
        lz = A.Lanczos(op, "f32", max_steps=nl + 2)
        lz.start(seed=1)
        ev.sync()
        e2, e3 = ev.event(), ev.event()
        ev.record(e2)
        lz.steps(nl)
        ev.record(e3)
        out["lanczos_ms_per_step"] = round(ev.elapsed_ms(e2, e3) / nl, 5)
        lz.close()
    if want("eigen"):
        out["eigen"] = eigen_leg(model, L, ev, op)
    if want("j1j2"):
        out["j1j2"] = j1j2_leg(ev)
    if want("bloch"):
        out["bloch"] = bloch_leg(ev, L)
    if want("batch"):
        out["batch"] = batch_leg(args.batch_size, args.seq_sample, graphs=args.graphs)
    if want("serve"):
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        out["serve"] = serve_leg(args.serve_requests, rates, graphs=args.graphs)
    if want("sweep"):
        out["sweep"] = sweep_leg()
    if want("cpu"):
        out["cpu_baseline"] = cpu_baseline(L, min(args.cpu_rows, D))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
