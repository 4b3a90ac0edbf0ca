"""Continuous-batching serving loop over the batch engine (scn_batch_* of include/scone_gpu.h).

Requests are fixed-source solves.  A request is admitted as soon as a slot and enough of the HBM
budget are free (scn_batch_add refuses it otherwise and it waits, first come first served).
Every engine cycle advances all resident solves by `window` batches each (one launch transports
them together and closes them in order), so the work of requests that arrive while others run
overlaps on the device, and a request's latency is a few cycles rather than one cycle per batch.
With `adapt`, a solve with a stop bin transports 2 batches in its first cycle and afterwards only
the batches its stop bin's trend still asks for, so little work is spent on batches past the stop.
A solve leaves when its batches are done or when the relative standard deviation of its stop bin
reaches the target; its slot and memory return to the pool at once and the next waiting request
takes them.

Each solve's histories depend only on its own seed and tables, so its result is the same alone
and in any mix (tallies to summation order).

A cycle ends when its last history does, and a fixed-source history includes every secondary of
its fission chain, transported one after another on the history's own random-number stream: in a
multiplying medium the longest chain of a cycle (thousands of collisions among millions of
histories) keeps one lane busy long after the rest of the device has drained.  `Pool` serves one
queue with several engines, each on its own HIP stream and host thread (engine calls release the
GIL), so the transport of one engine fills the device while another waits on its longest chain.
"""
from __future__ import annotations

import collections
import contextlib
import gc
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import tables as T
from .engine import Batch, Busy


@contextlib.contextmanager
def _no_cyclic_gc():
    """Pause the cyclic garbage collector for a serving run.  A full collection walks every
    resident request object (1.35 s with 3 million requests on one host core) and stalls admission
    and harvest for that long; the loop makes no reference cycles, so nothing accumulates."""
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


@dataclass
class Request:
    comp: object            # tables.Compiled
    src: object             # tables.CompiledSource
    pop: int
    n_batches: int
    seed: int
    stop_bin: int = -1
    stop_rel: float = 0.0
    record: bool = False
    buffer: int = 0         # depth of a history's buffer of secondaries (0: the package default, 50)
    window: int = 1         # batches transported together per engine cycle (closed in order)
    adapt: int = 0          # 1-100 with a stop bin: each cycle's window follows the stop bin's trend
    arrival: float = 0.0    # seconds (wall mode) or cycle index (cycle mode)
    # filled by the server
    admitted: float = -1.0
    done: float = -1.0
    slot: int = -1
    result: object = None
    bins: object = None
    hist_rng: object = None
    hist_coll: object = None
    extra: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class Server:
    def __init__(self, batch: Batch):
        self.B = batch
        self.cycles = 0
        self.busy_rejections = 0

    def _admit(self, waiting, running, now):
        while waiting:
            rq = waiting[0]
            try:
                rq.slot = self.B.add(rq.comp, rq.src, rq.pop, rq.n_batches, rq.seed, stop_bin=rq.stop_bin,
                                     stop_rel=rq.stop_rel, record=rq.record, buffer=rq.buffer,
                                     window=rq.window, adapt=rq.adapt)
            except Busy:
                self.busy_rejections += 1
                return
            rq.admitted = now
            running[rq.slot] = rq
            waiting.popleft()

    def _harvest(self, running, now, keep):
        st = self.B.states()
        for slot in [s for s in running if st[s] == T.SLOT_DONE]:
            rq = running.pop(slot)
            r, bins, hr, hc = self.B.query(slot)
            rq.result = r
            rq.done = now()
            if keep:
                rq.bins, rq.hist_rng, rq.hist_coll = bins, hr, hc
            self.B.release(slot)

    def run_cycles(self, requests, keep=True):
        """Deterministic mode: request.arrival is the cycle before which it is submitted."""
        order = sorted(requests, key=lambda r: r.arrival)
        pending = collections.deque(order)
        waiting = collections.deque()
        running = {}
        c = 0
        while pending or waiting or running:
            while pending and pending[0].arrival <= c:
                waiting.append(pending.popleft())
            self._admit(waiting, running, c)
            if running:
                self.B.cycle(1)
                self.cycles += 1
            c += 1
            self._harvest(running, lambda: c, keep)
        return requests

    def run_wall(self, requests, keep=False, progress=None):
        """Wall-clock mode: request.arrival is seconds after the start.  Returns the elapsed time."""
        with _no_cyclic_gc():
            return self._run_wall(requests, keep, progress)

    def _run_wall(self, requests, keep, progress):
        order = sorted(requests, key=lambda r: r.arrival)
        pending = collections.deque(order)
        waiting = collections.deque()
        running = {}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        last = 0.0
        while pending or waiting or running:
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.popleft())
            self._admit(waiting, running, t)
            if running:
                self.B.cycle(1)
                self.cycles += 1
                self._harvest(running, now, keep)
            elif pending:
                time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
            if progress and now() - last > 20.0:
                last = now()
                progress(last, len(pending), len(waiting), len(running))
        return now()


class Pool:
    """Several batch engines serving one first-come-first-served queue in wall-clock mode.  Each
    engine has its own HIP stream and its own host thread; a waiting request goes to the first
    engine with a free slot and enough budget."""

    def __init__(self, batches):
        self.servers = [Server(b) for b in batches]

    @property
    def cycles(self):
        return sum(s.cycles for s in self.servers)

    @property
    def busy_rejections(self):
        return sum(s.busy_rejections for s in self.servers)

    def run_wall(self, requests, keep=False, progress=None):
        """request.arrival is seconds after the start.  Returns the elapsed time."""
        with _no_cyclic_gc():
            return self._run_wall(requests, keep, progress)

    def _run_wall(self, requests, keep, progress):
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        lock = threading.Lock()
        active = [0]          # requests admitted and not yet harvested, over all engines
        errors = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0

        def worker(srv):
            running = {}
            try:
                while True:
                    with lock:
                        t = now()
                        while pending and pending[0].arrival <= t:
                            waiting.append(pending.popleft())
                        n0 = len(running)
                        srv._admit(waiting, running, t)
                        active[0] += len(running) - n0
                        finished = not pending and not waiting and active[0] == 0
                        nxt = pending[0].arrival if pending else None
                    if finished or errors:
                        return
                    if running:
                        srv.B.cycle(1)
                        srv.cycles += 1
                        n0 = len(running)
                        srv._harvest(running, now, keep)
                        if len(running) != n0:
                            with lock:
                                active[0] -= n0 - len(running)
                    else:
                        time.sleep(0.0005 if nxt is None else max(0.0, min(0.002, nxt - now())))
            except BaseException as e:      # reported by the caller's thread
                errors.append(e)

        threads = [threading.Thread(target=worker, args=(s,), daemon=True) for s in self.servers]
        for th in threads:
            th.start()
        last = 0.0
        while any(th.is_alive() for th in threads):
            for th in threads:
                th.join(timeout=1.0)
            if progress and now() - last > 20.0:
                last = now()
                with lock:
                    progress(last, len(pending), len(waiting), active[0])
        if errors:
            raise errors[0]
        return now()


def latency_summary(requests):
    lat = np.array([r.latency for r in requests if r.done >= 0])
    if lat.size == 0:
        return {"n": 0}
    return {"n": int(lat.size), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}


def batched_workload_requests(transport="transportOperatorDT", n_problems=512, seed=11, stop_rel=0.01, window=10,
                              adapt=50, problems=None):
    """A factory of serving requests drawn from the batched fixed-source workload: each request is
    one of its `n_problems` problems (configs.batch_problem), picked at random, on a seed of its
    own, stopping at `stop_rel` relative sd of its total-flux bin (bin 0).  `problems`: the
    workload already compiled, as (comp, src, pop, cycles, ...) tuples.  Returns (make, rng):
    make(arrival) is the next request, rng the generator it draws from."""
    if problems is None:
        from . import configs
        problems = []
        for p in range(n_problems):
            inp, mg, pop, cycles = configs.batch_problem(p, transport=transport)
            comp = T.compile_problem(inp, mg)
            problems.append((comp, T.compile_source(inp["source"], comp), pop, cycles))
    rng = np.random.default_rng(seed)
    count = [0]

    def make(arrival):
        comp, src, pop, cycles = problems[int(rng.integers(len(problems)))][:4]
        count[0] += 1
        return Request(comp, src, pop, cycles, 700000 + count[0], stop_bin=0, stop_rel=stop_rel, arrival=arrival,
                       window=min(window, cycles), adapt=adapt)
    return make, rng


def main(argv=None):
    """Serving run on one GPU: capacity with `--saturate` requests waiting at t = 0, then a Poisson
    stream at `--load` of that capacity for `--seconds`; prints one JSON line (latency percentiles
    over the run and per minute, histories/s), progress on stderr every 20 s."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--load", type=float, default=0.9)
    ap.add_argument("--seconds", type=float, default=600.0)
    ap.add_argument("--saturate", type=int, default=4096)
    ap.add_argument("--slots", type=int, default=256)
    ap.add_argument("--budget-gib", type=float, default=8.0)
    args = ap.parse_args(argv)
    make, rng = batched_workload_requests()

    def engine():
        B = Batch(transport=T.TRANSPORT_DT, max_slots=args.slots, hbm_budget=int(args.budget_gib * (1 << 30)),
                  use_graph=1)
        return B, Server(B)

    B, srv = engine()
    reqs = [make(0.0) for _ in range(args.saturate)]
    dt = srv.run_wall(reqs)
    B.close()
    cap = args.saturate / dt
    cap_hist = sum(r.result.n_histories for r in reqs) / dt
    lam = args.load * cap
    arr = np.cumsum(rng.exponential(1.0 / lam, max(1, int(lam * args.seconds))))
    reqs = [make(float(a)) for a in arr]
    print(f"capacity {cap:.0f} requests/s; {len(reqs)} requests over {arr[-1]:.0f} s", file=sys.stderr, flush=True)
    B, srv = engine()
    el = srv.run_wall(reqs, progress=lambda t, p, w, r: print(f"t {t:.0f} s pending {p} waiting {w} running {r}",
                                                              file=sys.stderr, flush=True))
    B.close()
    lat = np.array([r.latency for r in reqs])
    out = latency_summary(reqs)
    minutes = [(arr >= 60 * m) & (arr < 60 * (m + 1)) for m in range(int(arr[-1] // 60))]
    out.update({"capacity_req_per_s": cap, "capacity_hist_per_s": cap_hist, "load": args.load,
                "offered_req_per_s": lam, "elapsed_s": el, "cycles": srv.cycles,
                "hist_per_s": sum(r.result.n_histories for r in reqs) / el, "busy_rejections": srv.busy_rejections,
                "early_stopped": int(sum(1 for r in reqs if r.result.n_batches < r.n_batches)),
                "p90_s": float(np.percentile(lat, 90)), "p999_s": float(np.percentile(lat, 99.9)),
                "p50_p99_per_minute_s": [[float(np.percentile(lat[m], 50)), float(np.percentile(lat[m], 99))]
                                         for m in minutes if m.any()]})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
