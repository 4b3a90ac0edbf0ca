"""Build libscone_amd.so (HIP, gfx950) in-tree with hipcc.

FP discipline: no contraction and no fast-math on host or device, so the engine's arithmetic is
the IEEE sequence the CPU oracle evaluates.  Packed FP32 instructions are disabled.  The two host
translation units and eighteen kernel objects (scn_kernels.hip once per engine, tracking mode,
geometry variant and, for the general-geometry kernels, nesting depth) compile in parallel, then link into one shared library.
"""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libscone_amd.so")
SRC = [os.path.join(HERE, "csrc", "scn_engine.hip"), os.path.join(HERE, "csrc", "scn_batch.hip")]
KERNELS = os.path.join(HERE, "csrc", "scn_kernels.hip")
# (engine, tracking mode, geometry variant, depth): 0 eigenvalue / 1 batch, ST 1 / DT 2 / HT 3,
# 0 lattice/pin/box (all depths in one object) / 1 general geometry (depth 3 or 8, one per object)
VARIANTS = [(e, m, 0, 0) for e in (0, 1) for m in (1, 2, 3)] + \
           [(e, m, 1, nl) for e in (0, 1) for m in (1, 2, 3) for nl in (3, 8)]
OBJ_DIR = os.path.join(HERE, "_build")
ARCH = os.environ.get("SCN_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
]


def _jobs():
    """(label, source, extra flags, object) of every translation unit."""
    jobs = []
    for src in SRC:
        base = os.path.splitext(os.path.basename(src))[0]
        jobs.append((base, src, [], os.path.join(OBJ_DIR, base + ".o")))
    for e, m, g, nl in VARIANTS:
        base = f"scn_kernels_e{e}_m{m}_g{g}" + (f"_nl{nl}" if g else "")
        defs = [f"-DSCN_KENGINE={e}", f"-DSCN_KMODE={m}", f"-DSCN_KGEO={g}"] + ([f"-DSCN_KNL={nl}"] if g else [])
        jobs.append((base, KERNELS, defs, os.path.join(OBJ_DIR, base + ".o")))
    # the slowest units first (general-geometry kernels, batch before eigenvalue, HT before ST/DT)
    jobs.sort(key=lambda j: (0 if "_g1" in j[0] else 1, 0 if "_e1" in j[0] else 1, 0 if "_m3" in j[0] else 1))
    return jobs


def build(verbose: bool = False, jobs: int = 0) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ_DIR, exist_ok=True)
    n_par = jobs or max(1, min(16, os.cpu_count() or 1))
    pending = _jobs()
    objs = [j[3] for j in pending]
    running, failed = [], False
    while pending or running:
        while pending and len(running) < n_par:
            label, src, defs, obj = pending.pop(0)
            cmd = [hipcc] + FLAGS + defs + ["-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            running.append((label, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        done = [r for r in running if r[1].poll() is not None]
        if not done:
            time.sleep(0.2)
            continue
        for r in done:
            running.remove(r)
            out, _ = r[1].communicate()
            if r[1].returncode != 0:
                sys.stderr.write(out)
                failed = True
    if failed:
        raise RuntimeError("hipcc failed compiling libscone_amd.so sources")
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC"] + objs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libscone_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
