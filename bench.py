#!/usr/bin/env python3
"""Throughput of the plain (one problem at a time) GPU engine on the 17x17 pin-lattice eigenvalue
problem: 7 groups, 1e6 histories per cycle, synthetic multigroup data from seed 2, FP64.

A step is one power-iteration cycle: transport of the whole bank (event kernels), fission-site
sort, canonical k-eff sums, tally close and reproducible population control, all on the device.
The W warmup steps are inactive cycles; the K timed steps are active cycles, which also score the
7 x 2 x 289 per-pin, per-group flux / fission-rate collision clerks.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "histories/s & p50/p99 solve latency, batched vs plain GPU, 1 MI355X"
HBM_PEAK_GBS = 8000.0
S_P = 96            # bytes of compulsory particle state per event (load + store counted twice below)
SCORE_B = 16        # FP64 read-modify-write per tally score
SITE_B = 64         # fission-site write


def algorithmic_bytes(d):
    """Event-based compulsory HBM traffic of the transport kernel (per event e:
    2*S_p + 16*n_scores(e) + 64*n_sites(e))."""
    events = d["n_segments"] + d["n_collisions"]
    return 2 * S_P * events + SCORE_B * d["n_scores"] + SITE_B * d["n_sites"]


def diff(a, b):
    return {k: b[k] - a[k] for k in a}


def run_gpu(transport, pop, warmup, steps, seed=2):
    from scone_amd import configs, tables as T
    from scone_amd.engine import Engine
    inp, mg = configs.pin_lattice_7g(pop=pop, inactive=warmup, active=steps, transport=transport)
    comp = T.compile_problem(inp, mg)
    eng = Engine(comp, pop)
    bank = eng.fission_source(pop, seed)
    eng.eigen_start(dict(pop=pop, inactive=warmup, active=steps, seed=seed), bank)
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
    timed_cycles = ct[warmup:warmup + steps]
    hist = d["n_histories"]
    kernel_s = d["transport_kernel_ms"] / 1e3
    launches = max(1, d["transport_launches"])
    bytes_alg = algorithmic_bytes(d)
    eng.close()
    return {
        "hist_per_s": hist / wall,
        "coll_per_s": d["n_collisions"] / wall,
        "ms_per_step": wall / steps * 1e3,
        "cycle_ms_p50": float(sorted(timed_cycles)[len(timed_cycles) // 2] * 1e3),
        "wall_s": wall,
        "histories": hist,
        "collisions": d["n_collisions"],
        "segments": d["n_segments"],
        "scores": d["n_scores"],
        "sites": d["n_sites"],
        "kernel_s": kernel_s,
        "launches": launches,
        "kernel_ms_per_launch": d["transport_kernel_ms"] / launches,
        "alg_bytes": bytes_alg,
        "achieved_gbs": bytes_alg / kernel_s / 1e9 if kernel_s > 0 else 0.0,
        "k_implicit": list(R.k_implicit),
        "k_last": float(kh[-1]),
    }


def cpu_baseline(transport, pop, seconds=10.0, seed=2):
    """The CPU oracle (OpenMP, history-based restatement) on the same workload: active cycles of
    the same bank size, repeated until ~`seconds` of wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import binding as O
    from scone_amd import configs, tables as T
    threads = min(16, os.cpu_count() or 1)
    inp, mg = configs.pin_lattice_7g(pop=pop, transport=transport)
    comp = T.compile_problem(inp, mg)
    bank = O.fission_source(comp, pop, seed)
    n_hist = 0
    cycles = 0
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
            "sample": f"{cycles} active cycle(s) of {pop} histories ({transport[-2:]}, 17x17 lattice, 7 groups, "
                      f"collision clerks on), OpenMP oracle, {dt:.1f} s"}


def measured_traffic(workload_key):
    path = os.path.join(ROOT, "profiles", "transport_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(workload_key)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--pop", type=int, default=1000000)
    ap.add_argument("--transport", default="transportOperatorDT",
                    choices=["transportOperatorDT", "transportOperatorST"])
    ap.add_argument("--compare", action="store_true", default=True,
                    help="also time the other tracking method")
    ap.add_argument("--no-compare", dest="compare", action="store_false")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported (this engine runs on a "
              f"single GPU; see the sharding design in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    print(f"[bench] plain engine, {args.transport}, pop {args.pop}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    main_r = run_gpu(args.transport, args.pop, args.warmup, args.steps)
    other = None
    if args.compare:
        other_t = "transportOperatorST" if args.transport == "transportOperatorDT" else "transportOperatorDT"
        print(f"[bench] compare {other_t}", file=sys.stderr, flush=True)
        other = run_gpu(other_t, args.pop, args.warmup, args.steps)
        other["transport"] = other_t
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.transport, args.pop, args.cpu_seconds)
    tag = args.transport[-2:]
    workload = f"pin_lattice_17x17_7g_pop{args.pop}_{tag}"
    traffic = measured_traffic(workload)
    achieved = main_r["achieved_gbs"]
    line = {
        "metric": METRIC,
        "value": main_r["hist_per_s"],
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded multigroup data, C5G7 UO2 pin map)",
        "config": {"workload": "plain MG eigenvalue, 17x17 square pin lattice (pitch 1.26, r 0.54), "
                               "reflective, 7 groups, 1e6 histories/cycle",
                   "pop": args.pop, "groups": 7, "transport": tag, "engine": "plain",
                   "tallies": "7 group clerks x (flux, fission) x 289 pins"},
        "collisions_per_s": main_r["coll_per_s"],
        "k_implicit": main_r["k_implicit"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "kernel": "k_transport", "kernel_ms_per_launch": main_r["kernel_ms_per_launch"],
                     "alg_bytes_per_launch": main_r["alg_bytes"] / main_r["launches"],
                     "launches": main_r["launches"]},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (main_r["hist_per_s"] / cpu["value"]) if cpu else None,
    }
    if other:
        line["compare"] = {other["transport"][-2:]: {k: other[k] for k in
                           ("hist_per_s", "coll_per_s", "ms_per_step", "achieved_gbs", "kernel_ms_per_launch",
                            "k_implicit")}}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
