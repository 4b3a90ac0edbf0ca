"""Physics packages on the MI355X engine: a SCONE input file in, SCONE's result file out.

`run_input(path)` reads the dictionary (the format SCONE reads), compiles the tables, runs
eigenPhysicsPackage%run (PhysicsPackages/eigenPhysicsPackage_class.f90:135-159) or
fixedSourcePhysicsPackage%run (fixedSourcePhysicsPackage_class.f90:135-147) on the GPU and writes
`outputFile` (default ./output) in asciiMATLAB through collectResults.  There is no CPU path: a
missing library or GPU raises.
"""
from __future__ import annotations

import os
import time

from . import dictio, output, tables as T
from .engine import Batch, Engine, secondary_bytes


def _seed(inp):
    if "seed" not in inp:
        raise ValueError("the input must give 'seed' (SCONE hashes the wall clock otherwise)")
    return int(inp["seed"])


def run_eigen(inp: dict, base_dir: str = "."):
    comp = T.compile_problem(inp, base_dir=base_dir)
    pop = int(inp["pop"])
    seed = _seed(inp)
    settings = dict(pop=pop, inactive=int(inp["inactive"]), active=int(inp["active"]), seed=seed,
                    keff0=float(inp.get("keff_0", 1.0)))
    c0 = time.process_time()
    eng = Engine(comp, pop)
    # eigenPhysicsPackage%generateInitialState: fissionSource over the geometry bounds, pRNG as given
    bank = eng.fission_source(pop, seed)
    t0 = time.perf_counter()
    R, kh, ba, bi = eng.eigen_run(settings, bank)
    transport = time.perf_counter() - t0
    eng.close()
    out = output.eigen_results(comp, inp, R, ba, cpu_time=time.process_time() - c0, transport_time=transport)
    return out, R, kh, ba


def run_fixed(inp: dict, base_dir: str = "."):
    comp = T.compile_problem(inp, base_dir=base_dir)
    src = T.compile_source(inp["source"], comp)
    pop, cycles, seed = int(inp["pop"]), int(inp["cycles"]), _seed(inp)
    c0 = time.process_time()
    buffer = int(inp.get("buffer", 50))
    # slot (queues, score memory, tables) plus the per-lane buffers of secondaries
    B = Batch(transport=comp.problem.transport, max_slots=1,
              hbm_budget=max(256 << 20, 4 * 400 * pop + 64 * int(comp.problem.n_bins) + (64 << 20))
              + secondary_bytes(comp.problem.transport, buffer))
    slot = B.add(comp, src, pop, cycles, seed, buffer=buffer)
    t0 = time.perf_counter()
    B.run_all()
    transport = time.perf_counter() - t0
    r, bins, _, _ = B.query(slot)
    B.close()
    out = output.fixed_results(comp, inp, bins, seed, pop, r.n_batches, cpu_time=time.process_time() - c0,
                               transport_time=transport)
    return out, r, bins


def run_input(path: str, output_path: str | None = None):
    """Run a SCONE input file on the GPU and write its result file; returns the file name."""
    inp = dictio.read(path)
    base = os.path.dirname(os.path.abspath(path))
    kind = inp.get("type")
    if kind == "eigenPhysicsPackage":
        out = run_eigen(inp, base)[0]
    elif kind == "fixedSourcePhysicsPackage":
        out = run_fixed(inp, base)[0]
    else:
        raise NotImplementedError(f"physics package {kind} is outside the engine's scope")
    return out.write(output_path or inp.get("outputFile", "./output"))
