"""Host logic of the serving loop (CPU): first-come-first-served admission under a slot and
memory limit, harvest of finished solves, latency bookkeeping.  The batch engine is replaced by a
stand-in with the same interface whose solves finish after a fixed number of cycles."""
import numpy as np

from scone_amd import tables as T
from scone_amd.engine import Busy
from scone_amd.serve import Request, Server, latency_summary


class FakeBatch:
    def __init__(self, slots, budget):
        self.slots = [None] * slots
        self.budget = budget
        self.used = 0
        self.log = []

    def add(self, comp, src, pop, n_batches, seed, stop_bin=-1, stop_rel=0.0, record=False, buffer=0, window=1,
            adapt=0):
        if self.used + pop > self.budget or None not in self.slots:
            raise Busy("full")
        s = self.slots.index(None)
        self.slots[s] = [pop, n_batches, 0, seed, max(1, window)]
        self.used += pop
        self.log.append(("add", seed))
        return s

    def cycle(self, n=1):
        for v in self.slots:
            if v is not None and v[2] < v[1]:
                v[2] = min(v[1], v[2] + v[4])
        return sum(1 for v in self.slots if v is not None and v[2] < v[1])

    def states(self):
        return np.array([T.SLOT_FREE if v is None else (T.SLOT_DONE if v[2] >= v[1] else T.SLOT_RUNNING)
                         for v in self.slots], dtype=np.int32)

    def query(self, s):
        r = T.FixedResult()
        v = self.slots[s]
        r.state, r.n_batches, r.n_histories = T.SLOT_DONE, v[2], v[0] * v[2]
        return r, np.zeros((1, 2)), None, None

    def release(self, s):
        self.used -= self.slots[s][0]
        self.slots[s] = None


def test_fifo_admission_and_latency():
    B = FakeBatch(slots=2, budget=100)
    reqs = [Request(None, None, pop, nb, seed=i, arrival=a) for i, (pop, nb, a) in
            enumerate([(60, 3, 0), (30, 2, 0), (50, 1, 0), (10, 1, 1), (40, 2, 5)])]
    Server(B).run_cycles(reqs)
    # request 2 (50) cannot join 0 (60) + 1 (30): it waits, and request 3 waits behind it (FIFO)
    order = [s for k, s in B.log if k == "add"]
    assert order == [0, 1, 2, 3, 4]
    assert all(r.done >= 0 for r in reqs)
    assert reqs[0].done == 3 and reqs[1].done == 2
    assert reqs[2].admitted >= reqs[1].done and reqs[3].admitted >= reqs[2].admitted
    assert [r.result.n_batches for r in reqs] == [3, 2, 1, 1, 2]
    s = latency_summary(reqs)
    assert s["n"] == 5 and s["p50_s"] <= s["p99_s"] <= s["max_s"]


def test_window_shortens_latency_in_cycles():
    B = FakeBatch(slots=4, budget=1000)
    reqs = [Request(None, None, 10, 10, seed=0, arrival=0, window=1),
            Request(None, None, 10, 10, seed=1, arrival=0, window=5),
            Request(None, None, 10, 10, seed=2, arrival=0, window=4)]
    Server(B).run_cycles(reqs)
    assert [r.done for r in reqs] == [10, 2, 3]
    assert [r.result.n_batches for r in reqs] == [10, 10, 10]


def test_pool_serves_every_request_once_across_engines():
    """Several engines, each on its own thread, drain one queue: every request is admitted by
    exactly one engine, in arrival order, and all of them finish with their full batch count."""
    from scone_amd.serve import Pool
    engines = [FakeBatch(slots=3, budget=100) for _ in range(3)]
    reqs = [Request(None, None, 20 + (i % 4) * 10, 1 + i % 5, seed=i, arrival=0.0005 * i, window=1 + i % 3)
            for i in range(60)]
    pool = Pool(engines)
    el = pool.run_wall(reqs)
    assert el > 0
    seeds = sorted(s for B in engines for k, s in B.log if k == "add")
    assert seeds == list(range(60))
    assert all(r.done >= r.arrival >= 0 for r in reqs)
    assert [r.result.n_batches for r in reqs] == [1 + i % 5 for i in range(60)]
    assert sum(1 for B in engines if any(k == "add" for k, _ in B.log)) > 1
    assert all(B.used == 0 and all(v is None for v in B.slots) for B in engines)
    assert pool.cycles >= 1


def test_pool_reports_an_engine_error():
    from scone_amd.serve import Pool

    class Failing(FakeBatch):
        def cycle(self, n=1):
            raise RuntimeError("engine failed")

    pool = Pool([Failing(slots=2, budget=100), Failing(slots=2, budget=100)])
    reqs = [Request(None, None, 10, 2, seed=i, arrival=0.0) for i in range(4)]
    try:
        pool.run_wall(reqs)
    except RuntimeError as e:
        assert "engine failed" in str(e)
    else:
        raise AssertionError("the engine's error was not raised")


def test_collector_paused_during_a_run_and_restored():
    """The serving loop runs with the cyclic collector paused (a full collection over millions of
    resident requests would stall it) and leaves it as it found it."""
    import gc

    class Watching(FakeBatch):
        seen = []

        def cycle(self, n=1):
            Watching.seen.append(gc.isenabled())
            return super().cycle(n)

    reqs = [Request(None, None, 10, 2, seed=i, arrival=0.0) for i in range(3)]
    assert gc.isenabled()
    Server(Watching(slots=2, budget=100)).run_wall(reqs)
    assert Watching.seen and not any(Watching.seen)
    assert gc.isenabled()
    gc.disable()
    try:
        Server(Watching(slots=2, budget=100)).run_wall([Request(None, None, 10, 1, seed=9, arrival=0.0)])
        assert not gc.isenabled()
    finally:
        gc.enable()


def test_batched_workload_requests_draw_from_the_problems():
    """Requests of the timed serving run: a problem of the batched workload each, drawn at random,
    on distinct seeds, stopping on its total-flux bin, with the window capped by its batches."""
    from scone_amd.serve import batched_workload_requests
    make, _ = batched_workload_requests(n_problems=3)
    reqs = [make(0.1 * i) for i in range(40)]
    assert len({r.seed for r in reqs}) == 40
    assert len({id(r.comp) for r in reqs}) == 3
    assert all(r.stop_bin == 0 and r.stop_rel == 0.01 and 1 <= r.window <= r.n_batches for r in reqs)
    assert [r.arrival for r in reqs] == [0.1 * i for i in range(40)]
    pre = [("c", "s", 7, 2, 99)]
    make2, _ = batched_workload_requests(problems=pre)
    r = make2(0.0)
    assert (r.comp, r.src, r.pop, r.n_batches, r.window) == ("c", "s", 7, 2, 2)
