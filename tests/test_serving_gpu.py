"""Serving controls of the batch engine through the C ABI (needs a GPU):
  - adaptive windows (scn_fixed_request.adapt): a solve with a stop bin transports 2 batches in its
    first cycle and then only what its stop bin's trend asks for; its histories, scores and stop
    batch equal the oracle's (and one batch per cycle), in fewer cycles than one batch per cycle;
  - scn_batch_set_window narrows a running solve's window (results unchanged) and rejects windows
    outside 1 .. the admitted one; scn_batch_bin returns getResult of one bin over the batches
    closed so far, equal to the oracle run for that many batches;
  - the per-lane buffers of secondaries come out of the HBM budget: a fissile solve that fits
    only without them is refused (Busy) while a non-fissile one of the same size is admitted, and
    a deeper buffer than the budget can hold is refused without disturbing the resident solves."""
import math

import numpy as np
import pytest

import binding as O
from helpers import assert_results_close
from scone_amd import configs, tables as T
from scone_amd.engine import Batch, Busy, EngineError, secondary_bytes

pytestmark = pytest.mark.gpu


def compiled(inp, mg):
    comp = T.compile_problem(inp, mg)
    return comp, T.compile_source(inp["source"], comp)


def oracle_ref(comp, src, pop, cycles, seed, stop_bin=-1, stop_rel=0.0, buffer=50):
    return O.fixed_run(comp, src, pop, cycles, seed, stop_bin=stop_bin, stop_rel=stop_rel, hist=True, threads=8,
                       buffer=buffer)


def adaptive_cycles(rel, target, n_batches, stop, window, adapt):
    """Cycles a solve takes under the documented window rule (first window 2, then adapt% of the
    batches still missing to N (rel_N / target)^2, at least 1, at most the window), given the stop
    bin's relative sd rel[N] after N batches."""
    N, cycles, w = 0, 0, min(2, window)
    while True:
        w = max(1, min(w, window, n_batches - N))
        N = min(N + w, stop)
        cycles += 1
        if N >= stop:
            return cycles
        w = max(1, min(window, math.ceil(adapt / 100 * (N * (rel[N] / target) ** 2 - N))))


def check_slot(B, slot, ref, label):
    res, done, h = ref
    r, bins, hr, hc = B.query(slot)
    assert r.state == T.SLOT_DONE, label
    assert r.n_batches == done, (label, r.n_batches, done)
    assert np.array_equal(hr, h["rng_end"]), label
    assert np.array_equal(hc, h["n_coll"]), label
    assert_results_close(res, bins, done, 1e-12, label)
    return bins


@pytest.mark.parametrize("adapt", [30, 50, 100])
def test_adaptive_window_same_results_fewer_cycles(adapt):
    print(f"case adaptive window {adapt}", flush=True)
    pop = 400
    inp, mg = configs.inf_medium_fixed(seed=3, nG=3, pop=pop, cycles=50, transport="transportOperatorDT")
    comp, src = compiled(inp, mg)
    full = oracle_ref(comp, src, pop, 50, 3)
    rel = full[0][0, 1] / full[0][0, 0]
    stop = oracle_ref(comp, src, pop, 50, 3, stop_bin=0, stop_rel=1.5 * rel)
    assert 10 <= stop[1] < 50
    never = oracle_ref(comp, src, pop, 12, 5, stop_bin=0, stop_rel=1e-9)   # target out of reach
    B = Batch(transport=comp.problem.transport, max_slots=4, hbm_budget=128 << 20)
    s0 = B.add(comp, src, pop, 50, 3, stop_bin=0, stop_rel=1.5 * rel, record=True, window=16, adapt=adapt)
    s1 = B.add(comp, src, pop, 12, 5, stop_bin=0, stop_rel=1e-9, record=True, window=16, adapt=adapt)
    s2 = B.add(comp, src, pop, 50, 3, record=True, window=7, adapt=adapt)   # no stop bin: adapt is ignored
    n, done = 0, {}
    while True:
        more = B.cycle(1)
        n += 1
        st = B.states()
        for s in (s0, s1, s2):
            if s not in done and st[s] == T.SLOT_DONE:
                done[s] = n
        if not more:
            break
    n0 = done[s0]
    # the cycle count follows the window rule on the oracle's trajectory of the stop bin
    traj = {N: oracle_ref(comp, src, pop, N, 3)[0][0] for N in range(2, stop[1])}
    want = adaptive_cycles({N: sd / m for N, (m, sd) in traj.items()}, 1.5 * rel, 50, stop[1], 16, adapt)
    assert n0 == want < stop[1] / 2, (n0, want, stop[1])
    a = check_slot(B, s0, stop, "adaptive, early stop")
    check_slot(B, s1, never, "adaptive, target out of reach")
    check_slot(B, s2, full, "no stop bin")
    assert done[s1] == 2             # 2 batches, then the rest of its 12 (target out of reach)
    assert done[s2] == -(-50 // 7)   # the solve without a stop bin kept its window of 7
    B.close()
    # one batch per cycle: the same bins bit for bit, in as many cycles as batches
    P = Batch(transport=comp.problem.transport, max_slots=1, hbm_budget=128 << 20)
    p = P.add(comp, src, pop, 50, 3, stop_bin=0, stop_rel=1.5 * rel, record=True, window=1)
    m = P.run_all()
    b = check_slot(P, p, stop, "one batch per cycle")
    P.close()
    assert m + 1 == stop[1]     # run_all counts the cycles after which solves were still running
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_adaptive_window_cycles_follow_the_trend():
    """The first cycle transports 2 batches; a solve whose target is out of reach then takes the
    rest of its window in one cycle, so it needs 2 cycles for 12 batches in a window of 10."""
    print("case adaptive trend", flush=True)
    inp, mg = configs.inf_medium_fixed(seed=3, nG=3, pop=300, cycles=12, transport="transportOperatorDT")
    comp, src = compiled(inp, mg)
    B = Batch(transport=comp.problem.transport, max_slots=1, hbm_budget=64 << 20)
    s = B.add(comp, src, 300, 12, 9, stop_bin=0, stop_rel=1e-9, window=10, adapt=100)
    seen = []
    while True:
        more = B.cycle(1)
        seen.append(B.bin_result(s, 0)[2])
        if not more:
            break
    B.close()
    assert seen == [2, 12], seen


def test_set_window_and_bin_result():
    print("case set_window / bin_result", flush=True)
    pop = 500
    inp, mg, _, _ = configs.lattice_request(1, transport="transportOperatorDT", scale=0.02)
    comp, src = compiled(inp, mg)
    B = Batch(transport=comp.problem.transport, max_slots=2,
              hbm_budget=(128 << 20) + secondary_bytes(comp.problem.transport, 1000))
    s = B.add(comp, src, pop, 9, 77, record=True, buffer=1000, window=4)
    assert B.bin_result(s, 0) == (0.0, 0.0, 0)
    B.cycle(1)
    m, sd, n = B.bin_result(s, 0)
    assert n == 4
    ref4 = oracle_ref(comp, src, pop, 4, 77, buffer=1000)
    assert_results_close(ref4[0][:1], np.array([[m, sd]]), 4, 1e-12, "bin 0 after 4 batches")
    with pytest.raises(EngineError):
        B.set_window(s, 5)          # wider than the window it was admitted with
    with pytest.raises(EngineError):
        B.set_window(s, 0)
    with pytest.raises(EngineError):
        B.bin_result(s, int(comp.problem.n_bins))
    B.set_window(s, 1)
    B.cycle(1)
    assert B.bin_result(s, 0)[2] == 5
    B.set_window(s, 3)
    B.run_all()
    check_slot(B, s, oracle_ref(comp, src, pop, 9, 77, buffer=1000), "windows 4, 1, 3")
    m, sd, n = B.bin_result(s, 0)
    assert n == 9
    B.close()


def test_budget_counts_secondary_buffers():
    print("case budget with secondaries", flush=True)
    t = T.TRANSPORT_DT
    stack = secondary_bytes(t, 50)
    assert stack > 0 and secondary_bytes(t, 1000) > stack
    fis_inp, fis_mg = configs.inf_medium_fixed_fissile(pop=2000, cycles=2, transport="transportOperatorDT")
    fis, fsrc = compiled(fis_inp, fis_mg)
    cap_inp, cap_mg = configs.inf_medium_fixed(pop=2000, cycles=2, transport="transportOperatorDT")
    cap, csrc = compiled(cap_inp, cap_mg)
    # a budget far smaller than the buffers: the non-fissile solve fits, the fissile one waits
    B = Batch(transport=t, max_slots=4, hbm_budget=64 << 20)
    with pytest.raises(Busy):
        B.add(fis, fsrc, 2000, 2, 11)
    s = B.add(cap, csrc, 2000, 2, 12, record=True)
    B.run_all()
    B.release(s)
    B.close()
    # with room for the default buffers: admitted; a solve that asks for a deeper buffer than the
    # budget holds waits, and the resident one is unaffected
    ref = oracle_ref(fis, fsrc, 2000, 2, 11)
    B = Batch(transport=t, max_slots=4, hbm_budget=(64 << 20) + stack)
    s = B.add(fis, fsrc, 2000, 2, 11, record=True)
    with pytest.raises(Busy):
        B.add(fis, fsrc, 2000, 2, 13, buffer=1000)
    B.run_all()
    check_slot(B, s, ref, "fissile next to a refused deeper buffer")
    B.release(s)
    # the buffers returned with the last fissile solve: a non-fissile solve can use the whole budget
    big = B.add(cap, csrc, 2000, 2, 12)
    B.run_all()
    B.release(big)
    B.close()
