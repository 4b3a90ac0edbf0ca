"""Fixed-point tallies (fixedPointTally 1: BIN sums as 128-bit fixed point in four 32-bit limb
words, scone_gpu.h fixed_point_tally) on the GPU, through the C ABI:
  - two eigenvalue runs give the same result bits, the host-driven and the graph-replayed cycles
    give the same result bits, and the results equal the oracle's FP64 sums to 1e-12 (the k
    history stays bit-exact: tallies do not touch the histories), for a score memory privatised
    in LDS and for one in global memory;
  - a batch of fixed-source problems (lattice slabs, 1-3D) gives the same result bits batched ten
    batches per launch and one problem at a time, and equals the oracle to 1e-12."""
import numpy as np
import pytest

import binding as O
from helpers import assert_results_close, problem, rel_close
from scone_amd import configs, tables as T
from scone_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,pop", [("inf_medium_2g", 3000), ("pin_lattice_7g", 2500)])
def test_eigen_fixed_point_results_are_reproducible(name, pop):
    print(f"case fixed point eigen {name}", flush=True)
    inp, mg = configs.CONFIGS[name](pop=pop, transport="transportOperatorDT")
    inp["fixedPointTally"] = 1
    comp = T.compile_problem(inp, mg)
    assert comp.problem.fixed_point_tally == 1
    S = dict(pop=pop, inactive=2, active=4, seed=7)
    bank = O.fission_source(comp, pop, O.rng_skip(7, 0))
    R0, kh0, ba0, _ = O.eigen_run(comp, S, bank)
    eng = Engine(comp, pop)
    runs = [eng.eigen_run(dict(S, cycle_graph=g), bank) for g in (0, 0, 1)]
    eng.close()
    for R, kh, ba, _ in runs:
        assert np.array_equal(kh0.view(np.uint64), kh.view(np.uint64))
        assert list(R0.k_implicit) == list(R.k_implicit)
        ok, worst = rel_close(ba0[:, 0], ba[:, 0], 1e-12)
        assert ok, worst
    b = [r[2] for r in runs]
    assert np.array_equal(b[0].view(np.uint64), b[1].view(np.uint64))
    assert np.array_equal(b[0].view(np.uint64), b[2].view(np.uint64))


def test_eigen_cycle_fixed_point_bins():
    print("case fixed point eigen_cycle", flush=True)
    inp, mg, comp = problem("pin_lattice_7g", pop=3000, transport="transportOperatorST")
    inp["fixedPointTally"] = 1
    comp = T.compile_problem(inp, mg)
    bank = O.fission_source(comp, 3000, 99)
    o_out, o_bins, o_ks, _ = O.eigen_cycle(comp, bank, 1234, 1.02)
    eng = Engine(comp, 3000)
    g1 = eng.eigen_cycle(bank, 1234, 1.02)
    g2 = eng.eigen_cycle(bank, 1234, 1.02)
    eng.close()
    ok, worst = rel_close(o_bins, g1[1], 1e-12)
    assert ok, worst
    assert np.array_equal(g1[1].view(np.uint64), g2[1].view(np.uint64))


def test_batch_fixed_point_batched_equals_plain_bit_for_bit():
    print("case fixed point batch", flush=True)
    probs = []
    for p in range(16):
        inp, mg, pop, cycles = configs.batch_problem(p * 31 + 7, transport="transportOperatorDT", scale=0.05)
        inp["fixedPointTally"] = 1
        comp = T.compile_problem(inp, mg)
        probs.append((comp, T.compile_source(inp["source"], comp), pop, cycles, inp["seed"]))
    refs = [O.fixed_run(c, s, pop, cyc, seed, threads=8)[0] for (c, s, pop, cyc, seed) in probs]
    B = Batch(transport=T.TRANSPORT_DT, max_slots=16, hbm_budget=2 << 30, tile=128, n_launch=1, use_graph=1)
    slots = [B.add(c, s, pop, cyc, seed, window=min(10, cyc)) for (c, s, pop, cyc, seed) in probs]
    B.run_all()
    batched = [B.query(s)[1] for s in slots]
    B.close()
    P = Batch(transport=T.TRANSPORT_DT, max_slots=1, hbm_budget=2 << 30, use_graph=0)
    for i, (c, s, pop, cyc, seed) in enumerate(probs):
        slot = P.add(c, s, pop, cyc, seed, window=1)
        P.run_all()
        plain = P.query(slot)[1]
        P.release(slot)
        assert np.array_equal(plain.view(np.uint64), batched[i].view(np.uint64)), f"problem {i}"
        assert_results_close(refs[i], plain, cyc, 1e-12, f"problem {i} vs oracle")
    P.close()
