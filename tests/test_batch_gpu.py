"""Parity of the batch (fixed-source, resident slots) engine with the oracle, through the C ABI
(needs a GPU).

Contract: every history bit-exact (its final RNG state and its collision count, which fix the
whole event sequence's random-number consumption); score results within 1e-12 relative (FP64
atomics change only the summation order); a solve gives the same histories alone and in a mix
of other solves, with or without the hipGraph replay, and after slot memory is reused."""
import numpy as np
import pytest

import binding as O
from helpers import assert_results_close
from scone_amd import configs, tables as T
from scone_amd.engine import secondary_bytes, Batch, Busy, EngineError

pytestmark = pytest.mark.gpu


def compiled(inp, mg):
    comp = T.compile_problem(inp, mg)
    return comp, T.compile_source(inp["source"], comp)


def oracle_ref(comp, src, pop, cycles, seed, stop_bin=-1, stop_rel=0.0, buffer=50):
    res, done, h = O.fixed_run(comp, src, pop, cycles, seed, stop_bin=stop_bin, stop_rel=stop_rel, hist=True,
                               threads=8, buffer=buffer)
    return res, done, h


def check_slot(B, slot, ref, label, counts=True):
    res, done, h = ref
    r, bins, hr, hc = B.query(slot)
    assert r.state == T.SLOT_DONE, label
    assert r.n_batches == done, (label, r.n_batches, done)
    bad = np.nonzero(hr != h["rng_end"])[0]
    assert len(bad) == 0, f"{label}: {len(bad)} histories end on another RNG state, first {bad[:5]}"
    assert np.array_equal(hc, h["n_coll"]), label
    assert_results_close(res, bins, done, 1e-12, label)
    if counts:   # (batches transported past an early stop count too)
        assert r.n_collisions == int(h["n_coll"].sum()), label


CASES = [
    ("inf-ST", lambda: configs.inf_medium_fixed(pop=2000, cycles=5, transport="transportOperatorST"), 2000, 5),
    ("inf-DT", lambda: configs.inf_medium_fixed(pop=2000, cycles=5, transport="transportOperatorDT"), 2000, 5),
    ("slab-ST", lambda: configs.slab_fixed(pop=1500, cycles=4, transport="transportOperatorST"), 1500, 4),
    ("slab-DT", lambda: configs.slab_fixed(pop=1500, cycles=4, transport="transportOperatorDT"), 1500, 4),
    ("fissile-ST", lambda: configs.inf_medium_fixed_fissile(pop=1500, cycles=4, transport="transportOperatorST"),
     1500, 4),
    ("fissile-DT", lambda: configs.inf_medium_fixed_fissile(pop=1500, cycles=4, transport="transportOperatorDT"),
     1500, 4),
    ("slab-track-ST", lambda: configs.slab_fixed(pop=1500, cycles=4, transport="transportOperatorST", track=True),
     1500, 4),
    ("slab-track-DT", lambda: configs.slab_fixed(pop=1500, cycles=4, transport="transportOperatorDT", track=True),
     1500, 4),
    ("slab-HT", lambda: configs.slab_fixed(pop=1500, cycles=4, transport="transportOperatorHT", track=True),
     1500, 4),
    ("fissile-HT", lambda: configs.inf_medium_fixed_fissile(pop=1500, cycles=4, transport="transportOperatorHT"),
     1500, 4),
]


@pytest.mark.parametrize("label,make,pop,cycles", CASES, ids=[c[0] for c in CASES])
def test_single_solve_bit_exact(label, make, pop, cycles):
    print(f"case {label}", flush=True)
    inp, mg = make()
    comp, src = compiled(inp, mg)
    ref = oracle_ref(comp, src, pop, cycles, inp["seed"])
    B = Batch(transport=comp.problem.transport, max_slots=4,
              hbm_budget=(256 << 20) + secondary_bytes(comp.problem.transport))
    s = B.add(comp, src, pop, cycles, inp["seed"], record=True)
    B.run_all()
    check_slot(B, s, ref, label)
    B.close()


def _mix(n, scale, transport):
    out = []
    for p in range(n):
        inp, mg, pop, cycles = configs.batch_problem(p * 29 + 3, transport=transport, scale=scale)
        comp, src = compiled(inp, mg)
        out.append((comp, src, pop, cycles, inp["seed"]))
    return out


@pytest.mark.parametrize("transport,use_graph", [("transportOperatorDT", 1), ("transportOperatorST", 0)])
def test_mixed_batch_matches_oracle_and_alone(transport, use_graph):
    print(f"case mix {transport} graph={use_graph}", flush=True)
    probs = _mix(12, 0.004, transport)
    tcode = probs[0][0].problem.transport
    refs = [oracle_ref(c, s, pop, cyc, seed) for (c, s, pop, cyc, seed) in probs]
    B = Batch(transport=tcode, max_slots=16, hbm_budget=2 << 30, tile=128, k0=4, n_launch=3, use_graph=use_graph)
    slots = [B.add(c, s, pop, cyc, seed, record=True) for (c, s, pop, cyc, seed) in probs]
    B.run_all()
    for i, (slot, ref) in enumerate(zip(slots, refs)):
        check_slot(B, slot, ref, f"mix problem {i}")
    st = B.stats()
    assert st.n_histories == sum(pop * cyc for (_, _, pop, cyc, _) in probs)
    B.close()
    # the same solve alone in its own engine gives the same bins bit for bit where the mix did
    c, s, pop, cyc, seed = probs[5]
    A = Batch(transport=tcode, max_slots=1, hbm_budget=512 << 20, use_graph=1 - use_graph)
    a = A.add(c, s, pop, cyc, seed, record=True)
    A.run_all()
    check_slot(A, a, refs[5], "alone")
    A.close()


def test_early_stop_and_slot_reuse():
    print("case early stop / reuse", flush=True)
    pop = 400
    inp, mg = configs.inf_medium_fixed(seed=3, nG=3, pop=pop, cycles=50, transport="transportOperatorDT")
    comp, src = compiled(inp, mg)
    full = oracle_ref(comp, src, pop, 50, 3)
    rel = full[0][0, 1] / full[0][0, 0]
    ref = oracle_ref(comp, src, pop, 50, 3, stop_bin=0, stop_rel=2.0 * rel)
    assert 2 <= ref[1] < 50
    B = Batch(transport=comp.problem.transport, max_slots=2, hbm_budget=64 << 20)
    s0 = B.add(comp, src, pop, 50, 3, stop_bin=0, stop_rel=2.0 * rel, record=True)
    inp2, mg2 = configs.slab_fixed(pop=700, cycles=3, transport="transportOperatorDT")
    comp2, src2 = compiled(inp2, mg2)
    ref2 = oracle_ref(comp2, src2, 700, 3, inp2["seed"])
    s1 = B.add(comp2, src2, 700, 3, inp2["seed"], record=True)
    with pytest.raises(Busy):
        B.add(comp2, src2, 700, 3, inp2["seed"])       # both slots taken
    B.run_all()
    check_slot(B, s0, ref, "early stop")
    check_slot(B, s1, ref2, "slab next to it")
    B.release(s0)
    B.release(s1)
    s2 = B.add(comp2, src2, 700, 3, inp2["seed"], record=True)   # reuses the freed memory
    B.run_all()
    check_slot(B, s2, ref2, "slab in reused slot")
    B.close()


@pytest.mark.parametrize("window", [2, 5, 16])
def test_batch_window_same_results_and_stop_point(window):
    """Several batches of a solve transported in one cycle and closed in order: the same histories,
    scores and early-stop batch as one batch per cycle, next to solves with other windows."""
    print(f"case window {window}", flush=True)
    pop = 400
    inp, mg = configs.inf_medium_fixed(seed=3, nG=3, pop=pop, cycles=50, transport="transportOperatorDT")
    comp, src = compiled(inp, mg)
    full = oracle_ref(comp, src, pop, 50, 3)
    rel = full[0][0, 1] / full[0][0, 0]
    ref = oracle_ref(comp, src, pop, 50, 3, stop_bin=0, stop_rel=2.0 * rel)
    assert 2 <= ref[1] < 50
    inp2, mg2 = configs.slab_fixed(pop=700, cycles=7, transport="transportOperatorDT", track=True)
    comp2, src2 = compiled(inp2, mg2)
    ref2 = oracle_ref(comp2, src2, 700, 7, inp2["seed"])
    B = Batch(transport=comp.problem.transport, max_slots=4, hbm_budget=128 << 20)
    s0 = B.add(comp, src, pop, 50, 3, stop_bin=0, stop_rel=2.0 * rel, record=True, window=window)
    s1 = B.add(comp2, src2, 700, 7, inp2["seed"], record=True, window=3)
    s2 = B.add(comp, src, pop, 50, 3, record=True, window=window)
    n = 0
    while B.cycle(1):
        n += 1
    check_slot(B, s0, ref, "early stop", counts=False)
    check_slot(B, s1, ref2, "slab, window 3")
    check_slot(B, s2, full, "full run")
    assert n + 1 == max(-(-50 // window), 3)
    B.close()
    with pytest.raises(EngineError, match="n_launch"):
        C = Batch(transport=comp.problem.transport, max_slots=1, hbm_budget=64 << 20, n_launch=2)
        try:
            C.add(comp, src, pop, 5, 3, window=2)
        finally:
            C.close()


def test_fissile_lattice_with_secondaries_and_buffer_overflow():
    """Fission chains of the 17x17 lattice (k_inf 0.87) followed as secondaries on the history's
    stream; a buffer too small for them is the package's 'Run out of space' error."""
    print("case lattice secondaries", flush=True)
    inp, mg, pop, cycles = configs.lattice_request(3, scale=0.02)
    comp, src = compiled(inp, mg)
    ref = oracle_ref(comp, src, pop, 3, inp["seed"], buffer=1000)
    assert ref[2]["n_coll"].max() > 50
    B = Batch(transport=comp.problem.transport, max_slots=2,
              hbm_budget=(256 << 20) + secondary_bytes(comp.problem.transport, 1000), n_launch=3, k0=4)
    s = B.add(comp, src, pop, 3, inp["seed"], record=True, buffer=1000)
    B.run_all()
    check_slot(B, s, ref, "lattice")
    B.close()
    with pytest.raises(RuntimeError, match="Run out of space"):
        O.fixed_run(comp, src, pop, 1, inp["seed"], buffer=1)
    B = Batch(transport=comp.problem.transport, max_slots=1,
              hbm_budget=(256 << 20) + secondary_bytes(comp.problem.transport, 1))
    B.add(comp, src, pop, 1, inp["seed"], buffer=1)
    with pytest.raises(EngineError, match="Run out of space"):
        B.run_all()
    B.close()


def test_budget_admission():
    print("case budget", flush=True)
    inp, mg = configs.inf_medium_fixed(pop=100000, cycles=1, transport="transportOperatorDT")
    comp, src = compiled(inp, mg)
    B = Batch(transport=comp.problem.transport, max_slots=8, hbm_budget=32 << 20)
    s = B.add(comp, src, 100000, 1, 3)
    with pytest.raises(Busy):
        B.add(comp, src, 100000, 1, 3)                  # a second 2 x 100000-history queue does not fit
    B.run_all()
    B.release(s)
    B.add(comp, src, 100000, 1, 3)
    B.close()


def test_serving_results_same_alone_and_in_mix():
    """Continuous batching: requests arriving at different cycles share the device with the ones
    already resident; each result equals the request run alone."""
    print("case serving mix", flush=True)
    from scone_amd.serve import Request, Server
    inp, mg, _, _ = configs.lattice_request(0, transport="transportOperatorDT", scale=0.005)
    comp, src = compiled(inp, mg)
    reqs = []
    for i in range(10):
        _, _, pop, cycles = configs.lattice_request(i, transport="transportOperatorDT", scale=0.005)
        reqs.append(Request(comp, src, pop, cycles, 500 + i, stop_bin=0, stop_rel=0.05, record=True, buffer=1000,
                            arrival=0 if i < 7 else i - 5, window=1 + 3 * (i % 2), adapt=50 * (i % 3 == 0)))
    stack = secondary_bytes(comp.problem.transport, 1000)
    B = Batch(transport=comp.problem.transport, max_slots=4, hbm_budget=(256 << 20) + stack)
    srv = Server(B)
    srv.run_cycles(reqs)
    assert srv.busy_rejections > 0          # more requests than slots: some had to wait
    B.close()
    for i, rq in enumerate(reqs):
        ref = oracle_ref(comp, src, rq.pop, rq.n_batches, rq.seed, stop_bin=0, stop_rel=0.05, buffer=1000)
        assert rq.result.n_batches == ref[1], i
        assert np.array_equal(rq.hist_rng, ref[2]["rng_end"]), i
        assert_results_close(ref[0], rq.bins, ref[1], 1e-12, f"request {i}")
        A = Batch(transport=comp.problem.transport, max_slots=1, hbm_budget=(64 << 20) + stack)
        a = A.add(comp, src, rq.pop, rq.n_batches, rq.seed, stop_bin=0, stop_rel=0.05, record=True, buffer=1000)
        A.run_all()
        r, bins, hr, hc = A.query(a)
        A.close()
        assert np.array_equal(hr, rq.hist_rng) and np.array_equal(hc, rq.hist_coll), i
        assert_results_close(bins, rq.bins, r.n_batches, 1e-12, f"request {i} alone")


@pytest.mark.parametrize("top,n_launch", [("transportOperatorST", 1), ("transportOperatorDT", 3)])
def test_weight_windows_bit_exact(tmp_path, top, n_launch):
    """neutronMGimp weight windows: split copies go to the history's buffer, roulette draws on the
    history's stream; with parking between launches the split count travels with the particle."""
    print(f"case weight windows {top} n_launch={n_launch}", flush=True)
    path = tmp_path / "ww"
    path.write_text(configs.ww_file_text([0.6, 0.8], [1.3, 2.0]))
    pop, cycles = 1500, 3
    inp, mg = configs.ww_medium(str(path), pop=pop, cycles=cycles, transport=top, max_split=20)
    comp, src = compiled(inp, mg)
    ref = oracle_ref(comp, src, pop, cycles, inp["seed"], buffer=inp["buffer"])
    B = Batch(transport=comp.problem.transport, max_slots=2,
              hbm_budget=(128 << 20) + secondary_bytes(comp.problem.transport, inp["buffer"]), n_launch=n_launch, k0=2)
    s = B.add(comp, src, pop, cycles, inp["seed"], record=True, buffer=inp["buffer"])
    B.run_all()
    check_slot(B, s, ref, f"weight windows {top}")
    B.close()
