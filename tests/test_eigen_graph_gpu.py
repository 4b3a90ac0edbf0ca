"""Device-resident power iteration (cycle_graph = 1): every cycle replayed from a captured hipGraph,
with the site order, sums, k estimates and normSize_Repr all on the device.  Against the host-driven
engine (cycle_graph = 0) and the C oracle through the C ABI: k history and the k estimates bit for
bit, tallies to 1e-12.  keff0 far from the true k makes the first cycles run the population control
through every regime (drop sites, whole copies plus duplicates)."""
import numpy as np
import pytest

import binding as O
from helpers import problem, rel_close
from scone_amd.engine import Engine

pytestmark = pytest.mark.gpu

CASES = [
    ("inf_medium_2g", "transportOperatorST", 1.0, 0),
    ("pin_lattice_7g", "transportOperatorDT", 1.0, 0),
    ("pin_lattice_7g", "transportOperatorDT", 3.0, 1),
    ("pin_lattice_7g", "transportOperatorST", 0.3, 0),
]


@pytest.mark.parametrize("name,top,keff0,itally", CASES,
                         ids=[f"{c[0]}-{c[1][-2:]}-k{c[2]}-t{c[3]}" for c in CASES])
def test_graph_cycles_equal_host_cycles_and_oracle(name, top, keff0, itally):
    print(f"case {name} {top} keff0 {keff0} inactive_tally {itally}", flush=True)
    pop = 2500
    inp, mg, comp = problem(name, pop=pop, transport=top)
    S = dict(pop=pop, inactive=4, active=5, seed=21, keff0=keff0, inactive_tally=itally)
    bank = O.fission_source(comp, pop, O.rng_skip(21, 0))
    R0, kh0, ba0, bi0 = O.eigen_run(comp, S, bank)
    eng = Engine(comp, pop)
    R1, kh1, ba1, bi1 = eng.eigen_run(dict(S, cycle_graph=0), bank)
    R2, kh2, ba2, bi2 = eng.eigen_run(dict(S, cycle_graph=1), bank)
    for kh in (kh1, kh2):
        assert np.array_equal(kh0.view(np.uint64), kh.view(np.uint64)), (kh0, kh)
    for R in (R1, R2):
        assert list(R0.k_analog) == list(R.k_analog)
        assert list(R0.k_implicit) == list(R.k_implicit)
        assert [list(x) for x in R0.imp] == [list(x) for x in R.imp]
    assert R2.n_histories == R1.n_histories == pop * 9
    ok, worst = rel_close(ba0[:, 0], ba2[:, 0], 1e-12)
    assert ok, worst
    if itally:
        ok, worst = rel_close(bi0[:, 0], bi2[:, 0], 1e-12)
        assert ok, worst
    assert np.all(eng.last_cycle_time > 0)
    eng.close()


@pytest.mark.parametrize("window", [1e-9, 1000.0])
def test_graph_threshold_search_paths(window):
    """normSize_Repr's threshold is found exactly whichever way the search goes: a window so
    narrow that the threshold always falls outside it (a pass over all draws), and one so wide
    that it holds more draws than the rank count takes (a radix select of the gathered draws).
    keff0 3 makes the first cycles drop most sites, then the population is near pop."""
    print(f"case threshold window {window}", flush=True)
    pop = 6000
    inp, mg, comp = problem("pin_lattice_7g", pop=pop, transport="transportOperatorDT")
    S = dict(pop=pop, inactive=4, active=3, seed=33, keff0=3.0)
    bank = O.fission_source(comp, pop, O.rng_skip(33, 0))
    R0, kh0, ba0, _ = O.eigen_run(comp, S, bank)
    eng = Engine(comp, pop)
    R, kh, ba, _ = eng.eigen_run(dict(S, cycle_graph=1, select_window=window), bank)
    eng.close()
    assert np.array_equal(kh0.view(np.uint64), kh.view(np.uint64)), (kh0, kh)
    assert list(R0.k_implicit) == list(R.k_implicit)
    ok, worst = rel_close(ba0[:, 0], ba[:, 0], 1e-12)
    assert ok, worst


def test_graph_cycles_split_over_steps():
    """start, step(2), step(rest), finish: the graph run is resumable and gives the one-call bits."""
    print("case split steps", flush=True)
    pop = 2000
    inp, mg, comp = problem("pin_lattice_7g", pop=pop, transport="transportOperatorDT")
    S = dict(pop=pop, inactive=3, active=3, seed=5, cycle_graph=1)
    bank = O.fission_source(comp, pop, O.rng_skip(5, 0))
    R0, kh0, _, _ = O.eigen_run(comp, S, bank)
    eng = Engine(comp, pop)
    eng.eigen_start(S, bank)
    assert eng.eigen_step(2) == 2
    assert eng.eigen_step(10) == 4
    R, kh, ct, ba, _ = eng.eigen_finish()
    assert np.array_equal(kh0.view(np.uint64), kh.view(np.uint64))
    assert list(R0.k_implicit) == list(R.k_implicit)
    eng.close()
