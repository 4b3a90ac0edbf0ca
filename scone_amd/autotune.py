"""Autotuner of the batch engine's launch shape.

The cycle of the batch engine has four free parameters (scn_batch_settings): the work tile
(histories a wave takes from the queue at once), K0 (flight/collision pairs per particle in the
first launch; K doubles per launch), the number of launches per cycle (the last runs every
survivor to completion) and the refill threshold (idle lanes a wave gathers before it starts new
histories); with one launch per cycle the requests' window (batches of every solve transported
per cycle) is a fifth.  Small tiles balance uneven problems better, large ones cut
atomics on the work counter; small K0 compacts the queues early (fewer idle lanes) at the cost of
more launches; a wider window spreads a cycle's tail (the wait for its longest histories) over
more batches.  The tuner times one cycle of a representative sample for each candidate and keeps
the fastest per batch.  Results do not depend on the choice: every candidate runs the same
histories and closes the same batches.
"""
from __future__ import annotations

import itertools
import time

from .engine import Batch

TILES = (64, 128, 256, 512, 1024)
K0S = (4, 16)
LAUNCHES = (1, 2)
REFILLS = (32,)


def candidates(tiles=TILES, k0s=K0S, launches=LAUNCHES, windows=(1,), refills=REFILLS):
    """(tile, k0, n_launch, window, refill); with one launch per cycle K0 plays no part, and windows
    of more than one batch need one launch per cycle."""
    out = []
    for tile, nl, rf in itertools.product(tiles, launches, refills):
        if nl == 1:
            out.extend((tile, 8, 1, w, rf) for w in windows)
        else:
            out.extend((tile, k0, nl, 1, rf) for k0 in k0s)
    return out


def tune(problems, transport, hbm_budget, max_slots=None, tiles=TILES, k0s=K0S, launches=LAUNCHES, cycles=1,
         windows=(1,), refills=REFILLS, log=None):
    """problems: list of (comp, src, pop, seed).  Returns (best settings dict with the window,
    table of timings per batch)."""
    max_slots = max_slots or max(1, len(problems))
    table = []
    best = None
    for tile, k0, nl, w, rf in candidates(tiles, k0s, launches, windows, refills):
        B = Batch(transport=transport, max_slots=max_slots, hbm_budget=hbm_budget, tile=tile, k0=k0, n_launch=nl,
                  use_graph=1, refill=rf)
        for (comp, src, pop, seed) in problems:
            B.add(comp, src, pop, (cycles + 1) * w, seed, window=w)
        B.cycle(1)            # first cycle records the graph
        t = time.perf_counter()
        B.cycle(cycles)
        dt = (time.perf_counter() - t) / (cycles * w)
        B.close()
        table.append({"tile": tile, "k0": k0, "n_launch": nl, "window": w, "refill": rf, "ms_per_batch": dt * 1e3})
        if log:
            log(f"[autotune] tile {tile:4d} k0 {k0:3d} launches {nl} window {w:2d} refill {rf:2d}: "
                f"{dt * 1e3:8.3f} ms/batch")
        if best is None or dt < best[0]:
            best = (dt, {"tile": tile, "k0": k0, "n_launch": nl, "window": w, "refill": rf})
    return best[1], table
