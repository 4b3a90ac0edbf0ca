#!/usr/bin/env python3
"""Histories/s and solve latency of the MI355X engine, batched vs plain, plus the serving loop.

Default workload (the batched fixed-source benchmark): 512 independent fixed-source problems
(1D-3D boxes with 1-8 material regions, 2-70 groups, 1e3-1e6 histories split into 10 batches, point
or volume sources, a 1e3-1e5-bin mesh flux tally each; synthetic seeded data), FP64, delta tracking.
A step is one batch of every resident problem.  An engine cycle transports WINDOW = 10 steps (every
problem's next 10 batches, its whole 10-batch job, in one launch, closed in order).  The W warmup
steps are rounded up to a multiple of the window and run first on the same problems; then exactly
K timed steps (K a multiple of the window; otherwise the window falls back to the largest of 5, 2, 1
dividing K, and the line says so).  Every problem therefore runs W' + K batches.
  batched  all 512 problems resident in one batch engine, cycles replayed from a hipGraph, tile and
           launch shape from the autotuner
  plain    the same engine holding one problem at a time (each problem's batches run alone)
Both run the identical histories (checked on a sample of problems: score results equal to 1e-12).
Also reported: the serving loop (continuous batching with HBM-budget admission and 1 % early stop)
at 50/75/100 % of measured capacity with p50/p99 latency, on two request streams: solves drawn
from the 512 batched-workload problems (a random problem on a seed of its own per request), and
the 17x17 pin-lattice stream (fixed-source solves of the fissile lattice, fission chains followed
as secondaries); the plain
eigenvalue engine on the 17x17 lattice (1e6 histories per cycle) and the 3D core (1e7 per cycle).

The CPU baseline is the C oracle on the host: every one of the 512 problems, `--cpu-batches`
batches each, problems run concurrently (one per worker thread, largest first) on all the cores
this process may use.

--workload eigen prints the eigenvalue line instead.  Prints one JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "histories/s & p50/p99 solve latency, batched vs plain GPU, 1 MI355X"
HBM_PEAK_GBS = 8000.0
S_P = 96            # bytes of compulsory particle state per event (loaded and stored: 2 * S_P)
SCORE_B = 16        # FP64 read-modify-write per tally score
SITE_B = 64         # fission-site write
WINDOW = 10         # batches of every problem per engine cycle: the 10-batch job of each problem


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(d):
    """Event-based compulsory HBM traffic of the transport kernel (per event e:
    2*S_p + 16*n_scores(e) + 64*n_sites(e))."""
    events = d["n_segments"] + d["n_collisions"]
    return 2 * S_P * events + SCORE_B * d["n_scores"] + SITE_B * d["n_sites"]


def results_agree(a, b, n_batches, rtol=1e-12):
    """(mean, sd) results equal up to summation order: means within rtol, each sd within rtol
    times the condition number of sd = sqrt(CSUM2/(N(N-1)) - mean^2/(N-1))."""
    a = np.asarray(a, dtype=float).reshape(-1, 2)
    b = np.asarray(b, dtype=float).reshape(-1, 2)
    scale = np.maximum(np.abs(a[:, 0]), np.abs(b[:, 0]))
    if np.any(np.abs(a[:, 0] - b[:, 0]) > rtol * scale):
        return False
    m, s = a[:, 0], a[:, 1]
    N = n_batches
    with np.errstate(divide="ignore", invalid="ignore"):
        cond = np.where(s > 0, 0.5 * (1.0 + 3.0 * m * m / ((N - 1) * s * s)), 1.0)
    tol = rtol * np.maximum(cond, 1.0) * np.maximum(np.abs(s), np.abs(b[:, 1]))
    return bool(np.all(np.abs(s - b[:, 1]) <= tol + 1e-300))


def stats_dict(s):
    return {k: getattr(s, k) for k, _ in s._fields_}


def diff(a, b):
    return {k: b[k] - a[k] for k in a}


def measured_traffic(workload_key):
    """HBM bytes per launch of the workload's transport kernel from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes summarised in profiles/transport_traffic.json (None if not profiled), and
    the summary record itself."""
    path = os.path.join(ROOT, "profiles", "transport_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload_key)
    except (OSError, ValueError):
        rec = None
    return (rec["bytes_per_launch"] if rec else None), rec


# ---------------------------------------------------------------------------------------------
# batched fixed-source workload
# ---------------------------------------------------------------------------------------------
def batch_problems(n, transport, scale=1.0):
    from scone_amd import configs, tables as T
    out = []
    for p in range(n):
        inp, mg, pop, cycles = configs.batch_problem(p, transport=transport, scale=scale)
        comp = T.compile_problem(inp, mg)
        src = T.compile_source(inp["source"], comp)
        out.append((comp, src, pop, cycles, inp["seed"]))
    return out


def budget_for(problems, n_batches, slack=1.25, window=1):
    """Device bytes the problems need resident at once (queues, score memory with a BIN copy per
    batch of the window, tables)."""
    tot = 0
    for (comp, _, pop, _, _) in problems:
        depth = 3
        live = pop * (8 * (3 * depth + 7) + 4 * (depth + 7))
        tot += 2 * live + 8 * (window + 2) * int(comp.problem.n_bins) + (1 << 20)
    return int(tot * slack) + (256 << 20)


def _engine_settings(settings):
    return {k: v for k, v in settings.items() if k != "window"}


def run_batched(problems, transport, settings, warmup, steps, budget):
    """A step is one batch of every problem; an engine cycle carries `window` steps."""
    from scone_amd.engine import Batch
    w = settings.get("window", 1)
    assert warmup % w == 0 and steps % w == 0
    B = Batch(transport=transport, max_slots=len(problems), hbm_budget=budget, use_graph=1,
              **_engine_settings(settings))
    t = time.perf_counter()
    slots = [B.add(c, s, pop, warmup + steps, seed, window=w) for (c, s, pop, _, seed) in problems]
    t_admit = time.perf_counter() - t
    t_start = time.perf_counter()
    if warmup:
        B.cycle(warmup // w)
    s0 = stats_dict(B.stats())
    t0 = time.perf_counter()
    cyc_t = []
    for _ in range(steps // w):
        a = time.perf_counter()
        B.cycle(1)
        cyc_t.append(time.perf_counter() - a)
    t1 = time.perf_counter()
    s1 = stats_dict(B.stats())
    st = B.states()
    assert all(st[s] == 2 for s in slots), "batched run: some problems did not finish"
    results = {}
    for i in SAMPLE:
        if i < len(slots):
            results[i] = B.query(slots[i])[1]
    d = diff(s0, s1)
    B.close()
    wall = t1 - t0
    return {"hist_per_s": d["n_histories"] / wall, "coll_per_s": d["n_collisions"] / wall,
            "ms_per_step": wall / steps * 1e3, "wall_s": wall, "histories": d["n_histories"],
            "collisions": d["n_collisions"], "admit_s": t_admit,
            "latency_s": t1 - t_start,                 # every problem completes with the last cycle
            "cycle_ms": [x * 1e3 for x in cyc_t]}, results


def run_batched_kernel_timing(problems, transport, settings, budget, cycles=2):
    """Same workload without the graph, HIP events around every transport launch."""
    from scone_amd.engine import Batch
    w = settings.get("window", 1)
    B = Batch(transport=transport, max_slots=len(problems), hbm_budget=budget, use_graph=0,
              **_engine_settings(settings))
    for (c, s, pop, _, seed) in problems:
        B.add(c, s, pop, (cycles + 1) * w, seed, window=w)
    B.cycle(1)
    s0 = stats_dict(B.stats())
    B.cycle(cycles)
    s1 = stats_dict(B.stats())
    B.close()
    d = diff(s0, s1)
    launches = max(1, d["transport_launches"])
    kernel_s = d["transport_kernel_ms"] / 1e3
    alg = algorithmic_bytes(d)
    return {"kernel_ms_per_launch": d["transport_kernel_ms"] / launches, "launches": launches,
            "alg_bytes_per_launch": alg / launches, "achieved_gbs": alg / kernel_s / 1e9 if kernel_s > 0 else 0.0,
            "kernel_share": kernel_s / max(1e-12, d["cycle_ms"] / 1e3)}


def run_plain(problems, transport, settings, warmup, steps):
    from scone_amd.engine import Batch
    B = Batch(transport=transport, max_slots=1, hbm_budget=budget_for(problems[:1], 1) + max(
        budget_for([p], 1, window=settings.get("window", 1)) for p in problems), use_graph=1,
        **_engine_settings(settings))
    busy = 0.0
    hist = 0
    lat = []
    results = {}
    for i, (c, s, pop, _, seed) in enumerate(problems):
        slot = B.add(c, s, pop, warmup + steps, seed, window=settings.get("window", 1))
        t = time.perf_counter()
        B.run_all()
        busy += time.perf_counter() - t
        lat.append(busy)                 # all submitted at t = 0, served one after the other
        hist += pop * (warmup + steps)
        if i in SAMPLE:
            results[i] = B.query(slot)[1]
        B.release(slot)
        if i % 128 == 127:
            log(f"plain: {i + 1} problems, {busy:.1f} s")
    B.close()
    lat = np.array(lat)
    return {"hist_per_s": hist / busy, "wall_s": busy, "histories": hist,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99))}, results


SAMPLE = (0, 1, 2, 3, 100, 257, 511)


def host_cpu():
    """Cores this process may run on (affinity, capped by OMP_NUM_THREADS when set), the machine's
    logical CPU count and the CPU model."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    cores = min(allowed, int(omp)) if omp.isdigit() and int(omp) > 0 else allowed
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, {"nproc": os.cpu_count(), "affinity": allowed, "omp_num_threads": omp or None, "model": model}


def cpu_batch_baseline(problems, n_batches):
    """The C oracle on every problem for n_batches batches (the per-batch history counts of the
    workload), problems run concurrently: one single-threaded solve per worker thread, largest
    first (ctypes releases the interpreter lock during each call)."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import binding as O
    cores, host = host_cpu()
    order = sorted(range(len(problems)), key=lambda i: -problems[i][2])

    def solve(i):
        c, s, pop, _, seed = problems[i]
        O.fixed_run(c, s, pop, n_batches, seed, threads=1)
        return pop * n_batches

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        n_hist = sum(ex.map(solve, order))
    dt = time.perf_counter() - t0
    return {"value": n_hist / dt, "unit": "histories/s", "cores": cores, "kind": "port",
            "sample": f"all {len(problems)} problems, {n_batches} batches each ({n_hist} histories), one problem "
                      f"per worker thread, C oracle, {dt:.1f} s", "host": host}


# ---------------------------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------------------------
def serving(transport, seconds_per_load, loads=(0.5, 0.75, 1.0), stop_rel=0.01, seed=7, window=10, adapt=50,
            engines=1, slots=256, saturate=1024):
    """Continuous batching of 17x17-lattice fixed-source requests (1e3-1e6 histories in 10 batches,
    fission chains followed as secondaries, early stop at 1 % relative sd of the total flux).
    Capacity: `saturate` requests waiting at t = 0, served by one engine with fixed windows of
    `window` batches and with adaptive windows (`adapt`), and, for `engines` > 1, by a pool of that
    many engines (own stream and host thread each, `slots` slots in all) with adaptive windows.
    Latency at offered loads of the last of these capacities, served the same way.  Each engine's HBM budget holds
    the per-lane buffers of fission secondaries (the requests' 'buffer' of 1000 entries) next to
    its slots."""
    from scone_amd import configs, tables as T
    from scone_amd.engine import Batch, secondary_bytes
    from scone_amd.serve import Pool, Request, Server, latency_summary
    inp, mg, _, _ = configs.lattice_request(0, transport=transport)
    comp = T.compile_problem(inp, mg)
    src = T.compile_source(inp["source"], comp)

    def request(i, arrival, ad):
        _, _, pop, cycles = configs.lattice_request(i, transport=transport)
        return Request(comp, src, pop, cycles, 1000 + i, stop_bin=0, stop_rel=stop_rel, arrival=arrival,
                       buffer=inp["buffer"], window=min(window, cycles), adapt=ad)

    stack = secondary_bytes(comp.problem.transport, inp["buffer"])
    budget = (2 << 30) + stack

    def make(n_eng):
        Bs = [Batch(transport=comp.problem.transport, max_slots=slots // n_eng, hbm_budget=budget, use_graph=1)
              for _ in range(n_eng)]
        return Bs, (Server(Bs[0]) if n_eng == 1 else Pool(Bs))

    out = {"window": window, "adapt": adapt, "engines": engines, "slots": slots,
           "secondary_buffer_bytes_per_engine": stack, "hbm_budget_per_engine": budget, "loads": {}}
    legs = [("one_engine_fixed_window", 1, 0), ("one_engine_adaptive", 1, adapt)]
    if engines > 1:
        legs.append(("pool_adaptive", engines, adapt))
    for key, n_eng, ad in legs:
        Bs, srv = make(n_eng)
        reqs = [request(100000 + i, 0.0, ad) for i in range(saturate)]
        dt = srv.run_wall(reqs)
        hist = sum(r.result.n_histories for r in reqs)
        for B in Bs:
            B.close()
        out[key] = {"capacity_hist_per_s": hist / dt, "capacity_req_per_s": len(reqs) / dt, "cycles": srv.cycles,
                    "histories": hist, "elapsed_s": dt, "engines": n_eng}
        log(f"serving capacity ({key}) {hist / dt:.3e} histories/s, {len(reqs) / dt:.1f} requests/s, "
            f"{srv.cycles} cycles, {dt:.2f} s")
    served = out[legs[-1][0]]
    out.update(capacity_hist_per_s=served["capacity_hist_per_s"], capacity_req_per_s=served["capacity_req_per_s"],
               mean_request_histories=served["histories"] / saturate)
    cap_req = out["capacity_req_per_s"]
    rng = np.random.default_rng(seed)
    k = 0
    for load in loads:
        lam = load * cap_req
        n = max(1, int(lam * seconds_per_load))
        arr = np.cumsum(rng.exponential(1.0 / lam, n))
        reqs = [request(k + i, float(a), adapt) for i, a in enumerate(arr)]
        k += n
        Bs, srv = make(engines)
        el = srv.run_wall(reqs, progress=lambda t, p, w, r: log(f"serving load {load}: {t:.0f} s, pending {p}, "
                                                                 f"waiting {w}, running {r}"))
        hist = sum(r.result.n_histories for r in reqs)
        summ = latency_summary(reqs)
        summ.update({"offered_req_per_s": lam, "hist_per_s": hist / el, "elapsed_s": el, "cycles": srv.cycles,
                     "busy_rejections": srv.busy_rejections,
                     "early_stopped": int(sum(1 for r in reqs if r.result.n_batches < r.n_batches))})
        out["loads"][str(load)] = summ
        for B in Bs:
            B.close()
        log(f"serving load {load}: p50 {summ['p50_s'] * 1e3:.1f} ms, p99 {summ['p99_s'] * 1e3:.1f} ms, "
            f"{summ['hist_per_s']:.3e} histories/s")
    return out


def serving_mix(problems, transport, seconds_per_load, loads=(0.5, 0.75, 1.0), stop_rel=0.01, seed=11, window=10,
                adapt=50, slots=256, saturate=4096, budget=8 << 30):
    """Continuous batching of requests drawn from the batched workload's problems: each request is
    one of the 512 problems (boxes, 1-8 regions, 2-70 groups, 1e3-1e6 histories in 10 batches, a
    mesh flux tally), picked at random, on a seed of its own, stopping early at `stop_rel` relative
    sd of its total-flux bin (bin 0).  Capacity: `saturate` requests waiting at t = 0, adaptive
    windows; latency at offered Poisson loads of that capacity, `seconds_per_load` each."""
    from scone_amd import tables as T
    from scone_amd.engine import Batch
    from scone_amd.serve import Server, batched_workload_requests, latency_summary
    request, rng = batched_workload_requests(transport, seed=seed, stop_rel=stop_rel, window=window, adapt=adapt,
                                             problems=problems)

    def make():
        B = Batch(transport=T.TRANSPORT_DT if transport == "transportOperatorDT" else T.TRANSPORT_ST,
                  max_slots=slots, hbm_budget=budget, use_graph=1)
        return B, Server(B)

    out = {"window": window, "adapt": adapt, "slots": slots, "hbm_budget": budget, "stop_rel": stop_rel, "loads": {}}
    B, srv = make()
    reqs = [request(0.0) for _ in range(saturate)]
    dt = srv.run_wall(reqs)
    hist = sum(r.result.n_histories for r in reqs)
    B.close()
    out.update(capacity_hist_per_s=hist / dt, capacity_req_per_s=saturate / dt, capacity_cycles=srv.cycles,
               mean_request_histories=hist / saturate)
    log(f"serving (batched-workload requests) capacity {hist / dt:.3e} histories/s, {saturate / dt:.1f} requests/s, "
        f"{srv.cycles} cycles, {dt:.2f} s")
    cap_req = saturate / dt
    for load in loads:
        lam = load * cap_req
        n = max(1, int(lam * seconds_per_load))
        arr = np.cumsum(rng.exponential(1.0 / lam, n))
        reqs = [request(float(a)) for a in arr]
        B, srv = make()
        el = srv.run_wall(reqs)
        hist = sum(r.result.n_histories for r in reqs)
        B.close()
        summ = latency_summary(reqs)
        summ.update({"offered_req_per_s": lam, "hist_per_s": hist / el, "elapsed_s": el, "cycles": srv.cycles,
                     "busy_rejections": srv.busy_rejections,
                     "early_stopped": int(sum(1 for r in reqs if r.result.n_batches < r.n_batches))})
        out["loads"][str(load)] = summ
        log(f"serving (batched-workload requests) load {load}: p50 {summ['p50_s'] * 1e3:.1f} ms, "
            f"p99 {summ['p99_s'] * 1e3:.1f} ms, {summ['hist_per_s']:.3e} histories/s")
    return out


# ---------------------------------------------------------------------------------------------
# eigenvalue (plain engine, 17x17 lattice)
# ---------------------------------------------------------------------------------------------
def run_core3d(transport, pop=10_000_000, inactive=2, active=3, seed=5, cycle_graph=0):
    """3D nested core (pin -> 17x17 assembly -> 15x15 core with a 20 cm reflector, 400 cm tall,
    70 groups, 200x200x50x70 = 1.4e8-bin mesh flux tally), 1e7 histories per cycle: a bounded
    sample of its 30 + 100 cycles (inactive cycles untimed, active cycles timed with tallies).
    cycle_graph = 1: the device-resident cycles replayed from hipGraphs (no host round trip)."""
    from scone_amd import configs, tables as T
    from scone_amd.engine import Engine
    inp, mg = configs.core_3d(pop=pop, transport=transport)
    comp = T.compile_problem(inp, mg)
    eng = Engine(comp, pop)
    bank = eng.fission_source(pop, seed)
    eng.eigen_start(dict(pop=pop, inactive=inactive, active=active, seed=seed, cycle_graph=cycle_graph), bank)
    eng.eigen_step(inactive)
    s0 = eng.stats()
    t0 = time.perf_counter()
    done = eng.eigen_step(active)
    wall = time.perf_counter() - t0
    s1 = eng.stats()
    R, kh, _, _, _ = eng.eigen_finish()
    eng.close()
    assert done == active
    d = diff(s0, s1)
    out = {"hist_per_s": d["n_histories"] / wall, "coll_per_s": d["n_collisions"] / wall,
           "ms_per_cycle": wall / active * 1e3, "bins": int(comp.problem.n_bins),
           "k_implicit": list(R.k_implicit), "k_history": [float(x) for x in kh], "cycles_timed": active, "pop": pop,
           "engine": "cycle_graph" if cycle_graph else "plain"}
    if not cycle_graph:
        out["kernel_ms_per_launch"] = d["transport_kernel_ms"] / max(1, d["transport_launches"])
        out["kernel_share_of_cycle"] = d["transport_kernel_ms"] / max(1e-9, wall * 1e3)
    return out


def run_eigen(transport, pop, warmup, steps, seed=2, cycle_graph=0):
    from scone_amd import configs, tables as T
    from scone_amd.engine import Engine
    inp, mg = configs.pin_lattice_7g(pop=pop, inactive=warmup, active=steps, transport=transport)
    comp = T.compile_problem(inp, mg)
    eng = Engine(comp, pop)
    bank = eng.fission_source(pop, seed)
    eng.eigen_start(dict(pop=pop, inactive=warmup, active=steps, seed=seed, cycle_graph=cycle_graph), bank)
    eng.eigen_step(warmup)
    s0 = eng.stats()
    t0 = time.perf_counter()
    done = eng.eigen_step(steps)
    t1 = time.perf_counter()
    s1 = eng.stats()
    R, kh, ct, ba, _ = eng.eigen_finish()
    assert done == steps
    d = diff(s0, s1)
    wall = t1 - t0
    kernel_s = d["transport_kernel_ms"] / 1e3
    launches = max(1, d["transport_launches"])
    bytes_alg = algorithmic_bytes(d)
    eng.close()
    return {"hist_per_s": d["n_histories"] / wall, "coll_per_s": d["n_collisions"] / wall,
            "ms_per_step": wall / steps * 1e3, "kernel_ms_per_launch": d["transport_kernel_ms"] / launches,
            "launches": launches, "alg_bytes_per_launch": bytes_alg / launches,
            "achieved_gbs": bytes_alg / kernel_s / 1e9 if kernel_s > 0 else 0.0,
            "k_implicit": list(R.k_implicit)}


def cpu_eigen_baseline(transport, pop, seconds=10.0, seed=2):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import binding as O
    from scone_amd import configs, tables as T
    threads, host = host_cpu()
    inp, mg = configs.pin_lattice_7g(pop=pop, transport=transport)
    comp = T.compile_problem(inp, mg)
    bank = O.fission_source(comp, pop, seed)
    n_hist = cycles = 0
    master = seed
    t0 = time.perf_counter()
    while True:
        out, bins, ks, _ = O.eigen_cycle(comp, bank, master, 1.0, threads=threads)
        n_hist += len(bank)
        cycles += 1
        master = O.rng_skip(master, 152917 * (pop + 2))
        bank = O.norm_size_repr(out, pop, master)
        if time.perf_counter() - t0 >= seconds or cycles >= 50:
            break
    dt = time.perf_counter() - t0
    return {"value": n_hist / dt, "unit": "histories/s", "cores": threads, "kind": "port",
            "sample": f"{cycles} cycle(s) of {pop} histories ({transport[-2:]}, 17x17 lattice, 7 groups, "
                      f"collision clerks on), OpenMP oracle, {dt:.1f} s", "host": host}


def eigen_line(args):
    log(f"eigen: plain engine, {args.transport}, pop {args.pop}, warmup {args.warmup}, steps {args.steps}")
    r = run_eigen(args.transport, args.pop, args.warmup, args.steps)
    cpu = None if args.no_cpu else cpu_eigen_baseline(args.transport, args.pop, args.cpu_seconds)
    tag = args.transport[-2:]
    return {
        "metric": METRIC, "value": r["hist_per_s"], "unit": "histories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded multigroup data, C5G7 UO2 pin map)",
        "config": {"workload": "plain MG eigenvalue, 17x17 square pin lattice (pitch 1.26, r 0.54), reflective, "
                               "7 groups, 1e6 histories/cycle", "pop": args.pop, "groups": 7, "transport": tag,
                   "engine": "plain", "tallies": "7 group clerks x (flux, fission) x 289 pins"},
        "collisions_per_s": r["coll_per_s"], "k_implicit": r["k_implicit"],
        "roofline": {"bound": "hbm", "achieved": r["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": r["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"pin_lattice_17x17_7g_pop{args.pop}_{tag}")[0],
                     "traffic_source": measured_traffic(f"pin_lattice_17x17_7g_pop{args.pop}_{tag}")[1],
                     "kernel": "k_transport", "kernel_ms_per_launch": r["kernel_ms_per_launch"],
                     "alg_bytes_per_launch": r["alg_bytes_per_launch"], "launches": r["launches"]},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (r["hist_per_s"] / cpu["value"]) if cpu else None,
    }


def batch_line(args):
    from scone_amd import autotune, tables as T
    transport = T.TRANSPORT_DT if args.transport == "transportOperatorDT" else T.TRANSPORT_ST
    log(f"batch: compiling {args.problems} problems")
    problems = batch_problems(args.problems, args.transport, args.scale)
    window = next(w for w in (WINDOW, 5, 2, 1) if args.steps % w == 0)
    warmup = -(-args.warmup // window) * window
    if window != WINDOW:
        log(f"batch: {args.steps} timed steps are not a multiple of {WINDOW}: window {window}")
    budget = budget_for(problems, warmup + args.steps, window=window)
    tot_hist = sum(p[2] for p in problems)
    log(f"batch: {tot_hist} histories per batch, device budget {budget / 2**30:.2f} GiB, window {window}, "
        f"warmup {warmup} batches")
    if args.tune:
        t = time.perf_counter()
        best, table = autotune.tune([(c, s, pop, seed) for (c, s, pop, _, seed) in problems], transport, budget,
                                    tiles=(64, 128, 256, 512, 1024), launches=(1,), windows=(window,), refills=(16, 32),
                                    log=log)
        tune_s = time.perf_counter() - t
    else:
        best, table, tune_s = {"tile": 128, "k0": 8, "n_launch": 1, "window": window, "refill": 32}, [], 0.0
    log(f"batch: tuned {best} in {tune_s:.1f} s")
    bat, res_b = run_batched(problems, transport, best, warmup, args.steps, budget)
    log(f"batch: batched {bat['hist_per_s']:.3e} histories/s, {bat['ms_per_step']:.2f} ms/step")
    kt = run_batched_kernel_timing(problems, transport, best, budget)
    log(f"batch: transport kernel {kt['kernel_ms_per_launch']:.3f} ms/launch, {kt['achieved_gbs']:.1f} GB/s")
    pla, res_p = (None, {})
    if not args.no_plain:
        # one problem at a time has far fewer histories in flight: the tile that suits it is smaller
        pla, res_p = run_plain(problems, transport, dict(best, tile=128), warmup, args.steps)
        log(f"batch: plain {pla['hist_per_s']:.3e} histories/s over {pla['wall_s']:.1f} s")
    identical = None
    if res_p:
        identical = all(results_agree(res_b[i], res_p[i], warmup + args.steps) for i in res_b)
    cpu = None if args.no_cpu else cpu_batch_baseline(problems, args.cpu_batches)
    serve = None if args.serve_seconds <= 0 else serving(args.transport, args.serve_seconds, window=args.window,
                                                                  engines=args.engines)
    serve_mix = None if args.serve_seconds <= 0 else serving_mix(problems, args.transport, args.serve_seconds,
                                                                         window=args.window)
    eig = None
    if args.eigen:
        log("eigen leg")
        e = run_eigen(args.transport, args.pop, 3, 10)
        eg = run_eigen(args.transport, args.pop, 3, 10, cycle_graph=1)
        eig = {k: e[k] for k in ("hist_per_s", "coll_per_s", "ms_per_step", "achieved_gbs", "kernel_ms_per_launch",
                                 "k_implicit")}
        eig["cycle_graph"] = {k: eg[k] for k in ("hist_per_s", "ms_per_step", "k_implicit")}
        eig["cycle_graph"]["same_k_as_plain"] = eg["k_implicit"] == e["k_implicit"]
        log(f"eigen 17x17: plain {e['hist_per_s']:.3e}, graph {eg['hist_per_s']:.3e} histories/s")
    core = None
    if args.c5:
        log("core 3D leg")
        core = run_core3d(args.transport)
        cg = run_core3d(args.transport, cycle_graph=1)
        core["cycle_graph"] = {k: cg[k] for k in ("hist_per_s", "ms_per_cycle", "k_implicit")}
        core["cycle_graph"]["same_k_history_as_plain"] = cg["k_history"] == core["k_history"]
        log(f"core 3D: plain {core['hist_per_s']:.3e} histories/s, {core['ms_per_cycle']:.1f} ms/cycle; graph "
            f"{cg['hist_per_s']:.3e}, {cg['ms_per_cycle']:.1f} ms/cycle")
    line = {
        "metric": METRIC, "value": bat["hist_per_s"], "unit": "histories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": bat["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded problems and multigroup data)",
        "config": {"workload": f"batched fixed source: {args.problems} problems, 1D-3D boxes, 1-8 regions, 2-70 "
                               f"groups, a batch of each is 1/10 of its 1e3-1e6 histories, point/volume sources, "
                               f"1e3-1e5-bin mesh flux tallies; {warmup} warmup + {args.steps} timed batches per "
                               f"problem, {window} batches of every problem per launch",
                   "problems": args.problems, "histories_per_step": tot_hist, "transport": args.transport[-2:],
                   "batches_per_problem": warmup + args.steps, "warmup_batches": warmup, "timed_batches": args.steps,
                   "engine": "batched (persistent launch, hipGraph cycles, autotuned)", "launch_shape": best},
        "collisions_per_s": bat["coll_per_s"],
        "latency": {"batched_p50_s": bat["latency_s"], "batched_p99_s": bat["latency_s"],
                    "plain_p50_s": pla["latency_p50_s"] if pla else None,
                    "plain_p99_s": pla["latency_p99_s"] if pla else None,
                    "note": "all solves submitted at t=0; batched: every solve closes its last batch in the "
                            "same final cycle, so p50 = p99 = the whole run"},
        "plain": {"hist_per_s": pla["hist_per_s"], "wall_s": pla["wall_s"]} if pla else None,
        "batched_over_plain": (bat["hist_per_s"] / pla["hist_per_s"]) if pla else None,
        "identical_results_sample": identical,
        "roofline": {"bound": "hbm", "achieved": kt["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kt["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"batch512_{args.transport[-2:]}_w{best.get('window', 1)}")[0],
                     "traffic_source": measured_traffic(f"batch512_{args.transport[-2:]}_w{best.get('window', 1)}")[1],
                     "kernel": "k_btransport", "kernel_ms_per_launch": kt["kernel_ms_per_launch"],
                     "alg_bytes_per_launch": kt["alg_bytes_per_launch"], "launches": kt["launches"],
                     "kernel_share_of_cycle": kt["kernel_share"]},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (bat["hist_per_s"] / cpu["value"]) if cpu else None,
        "autotune": {"seconds": tune_s, "candidates": len(table)},
        "serving": serve,
        "serving_batched_requests": serve_mix,
        "eigen_17x17_1e6": eig,
        "eigen_core3d_70g_1e7": core,
    }
    return line


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="batch", choices=["batch", "eigen"])
    ap.add_argument("--transport", default="transportOperatorDT", choices=["transportOperatorDT", "transportOperatorST"])
    ap.add_argument("--problems", type=int, default=512)
    ap.add_argument("--scale", type=float, default=1.0, help="multiplies every problem's history count")
    ap.add_argument("--pop", type=int, default=1000000, help="eigenvalue histories per cycle")
    ap.add_argument("--no-tune", dest="tune", action="store_false", default=True)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=15.0, help="offered-load duration per load level")
    ap.add_argument("--window", type=int, default=10, help="serving: batches of a request transported per cycle")
    ap.add_argument("--engines", type=int, default=1, help="serving: engines (streams, host threads) in a pool")
    ap.add_argument("--no-eigen", dest="eigen", action="store_false", default=True)
    ap.add_argument("--no-c5", dest="c5", action="store_false", default=True)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="eigenvalue CPU baseline: seconds of cycles")
    ap.add_argument("--cpu-batches", type=int, default=10, help="batched CPU baseline: batches of every problem")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (the engine runs on a single "
              f"GPU; sharding the bank over GPUs is a design, not built)", file=sys.stderr)
        sys.exit(2)
    line = eigen_line(args) if args.workload == "eigen" else batch_line(args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
