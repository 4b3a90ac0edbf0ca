"""Workload driver and summariser for HBM-traffic counter runs of the transport kernels.

Run under rocprofv3 twice, once per counter (the two cannot share a pass with the traces):

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir_fetch> -- python3 -m scone_amd.profile_traffic run ...
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir_write> -- python3 -m scone_amd.profile_traffic run ...
    python3 -m scone_amd.profile_traffic summarize <dir_fetch> <dir_write> --key <workload key>

`run` first copies a 1 GiB device buffer with a plain device-to-device copy (hipMemcpy through
the HIP runtime the engine links, whose copy kernel is a separate dispatch), so the counters can be scaled: factor = counted KiB of that dispatch / its true KiB.  Then it runs the
workload without hipGraph replay, so every transport launch is a separately counted dispatch.
`summarize` averages FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel, divides them by
the copy's factor and writes bytes per launch of the transport kernel into
profiles/transport_traffic.json under the workload key.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COPY_MIB = 1024


def _calibration_copy():
    """One device-to-device copy of COPY_MIB MiB through the HIP runtime the engine links (its
    copy kernel is the dispatch the counters are scaled by)."""
    import ctypes as C
    from scone_amd.engine import load
    load()
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    nbytes = COPY_MIB << 20
    a, b = C.c_void_p(), C.c_void_p()
    if H.hipMalloc(C.byref(a), nbytes) or H.hipMalloc(C.byref(b), nbytes):
        raise SystemExit("hipMalloc for the calibration copy failed")
    H.hipMemset(a, 1, nbytes)
    if H.hipMemcpy(b, a, nbytes, 3) or H.hipDeviceSynchronize():   # hipMemcpyDeviceToDevice
        raise SystemExit("calibration copy failed")
    H.hipFree(a)
    H.hipFree(b)
    print(f"copy calibration: {COPY_MIB} MiB read + {COPY_MIB} MiB written", flush=True)


def run(args):
    sys.path.insert(0, ROOT)
    _calibration_copy()
    if args.workload == "batch512":
        import bench
        from scone_amd import tables as T
        from scone_amd.engine import Batch
        transport = T.TRANSPORT_DT if args.transport == "transportOperatorDT" else T.TRANSPORT_ST
        problems = bench.batch_problems(512, args.transport, 1.0)
        budget = bench.budget_for(problems, 2, window=args.window)
        B = Batch(transport=transport, max_slots=512, hbm_budget=budget, use_graph=0, tile=args.tile, k0=args.k0,
                  n_launch=args.n_launch, refill=args.refill)
        for (c, s, pop, _, seed) in problems:
            B.add(c, s, pop, args.cycles * args.window, seed, window=args.window)
        B.run_all()
        s = B.stats()
        st = {k: getattr(s, k) for k, _ in s._fields_}
        B.close()
    else:
        from scone_amd import configs, tables as T
        from scone_amd.engine import Engine
        if args.workload == "core3d":    # 70-group 3D core, 1e7 histories, 1.4e8-bin mesh tally
            pop, seed = 10_000_000, 5
            inp, mg = configs.core_3d(pop=pop, transport=args.transport)
        else:                            # 17x17 pin lattice, 7 groups, 1e6 histories
            pop, seed = 1_000_000, 2
            inp, mg = configs.pin_lattice_7g(pop=pop, transport=args.transport)
        comp = T.compile_problem(inp, mg)
        eng = Engine(comp, pop)
        bank = eng.fission_source(pop, seed)
        eng.eigen_start(dict(pop=pop, inactive=0, active=args.cycles, seed=seed, cycle_graph=args.cycle_graph), bank)
        eng.eigen_step(args.cycles)
        st = eng.stats()
        eng.eigen_finish()
        eng.close()
    print(json.dumps({"segments": st["n_segments"], "collisions": st["n_collisions"], "scores": st["n_scores"],
                      "sites": st["n_sites"], "launches": st["transport_launches"]}), flush=True)


def _read_counters(d):
    """{kernel name: [KiB per dispatch]} from a rocprofv3 csv counter-collection directory."""
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    per = {}
    for fn in files:
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                k = row.get("Kernel_Name") or row.get("Kernel-Name") or row.get("kernel_name")
                disp = row.get("Dispatch_Id") or row.get("Dispatch-Id") or row.get("dispatch_id")
                v = float(row.get("Counter_Value") or row.get("Counter-Value") or row.get("counter_value"))
                per.setdefault(k, {}).setdefault(disp, 0.0)
                per[k][disp] += v
    return {k: list(v.values()) for k, v in per.items()}


def summarize(args):
    fetch = _read_counters(args.fetch_dir)
    write = _read_counters(args.write_dir)

    def pick(d, *keys):
        for k in d:
            if all(s in k for s in keys):
                return k
        return None

    # the calibration copy is the largest dispatch of the runtime's copy kernel
    true_kib = COPY_MIB * 1024.0
    f_fetch = max(fetch[pick(fetch, "copyBuffer")]) / true_kib
    f_write = max(write[pick(write, "copyBuffer")]) / true_kib
    tk = pick(fetch, args.kernel)
    tw = pick(write, args.kernel)
    fk = sum(fetch[tk]) / len(fetch[tk])
    wk = sum(write[tw]) / len(write[tw])
    out = {"fetch_bytes_per_launch": fk * 1024.0 / f_fetch, "write_bytes_per_launch": wk * 1024.0 / f_write,
           "launches_counted": len(fetch[tk]), "copy_factor_fetch": f_fetch, "copy_factor_write": f_write,
           "kernel": tk}
    out["bytes_per_launch"] = out["fetch_bytes_per_launch"] + out["write_bytes_per_launch"]
    path = os.path.join(ROOT, "profiles", "transport_traffic.json")
    try:
        with open(path) as f:
            allv = json.load(f)
    except (OSError, ValueError):
        allv = {}
    allv[args.key] = out
    with open(path, "w") as f:
        json.dump(allv, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--workload", choices=["batch512", "eigen17", "core3d"], default="batch512")
    r.add_argument("--transport", default="transportOperatorDT")
    r.add_argument("--cycles", type=int, default=2)
    r.add_argument("--tile", type=int, default=0)
    r.add_argument("--k0", type=int, default=0)
    r.add_argument("--n-launch", type=int, default=0)
    r.add_argument("--window", type=int, default=1, help="batch512: batches of every problem per cycle")
    r.add_argument("--refill", type=int, default=0, help="batch512: idle lanes gathered before new histories start")
    r.add_argument("--cycle-graph", type=int, default=0, help="eigen: replay the device-resident cycles from hipGraphs")
    s = sub.add_parser("summarize")
    s.add_argument("fetch_dir")
    s.add_argument("write_dir")
    s.add_argument("--key", required=True)
    s.add_argument("--kernel", default="k_btransport")
    args = ap.parse_args()
    run(args) if args.cmd == "run" else summarize(args)


if __name__ == "__main__":
    main()
