"""Parity of the HIP engine with the CPU oracle, through the C ABI (needs a GPU).

Contract: per-history results bit-exact (next-cycle bank, every field; k-eff clerk scores, which
both sides sum in the same canonical order); collision-clerk BIN scores within 1e-12 relative
(the engine accumulates them with FP64 atomics, so only the summation order differs)."""
import numpy as np
import pytest

import binding as O
from helpers import assert_banks_identical, problem, rel_close
from scone_amd.engine import Engine

pytestmark = pytest.mark.gpu

CASES = [
    ("inf_medium_2g", dict(pop=4000), "transportOperatorST"),
    ("inf_medium_2g", dict(pop=4000), "transportOperatorDT"),
    ("pin_lattice_7g", dict(pop=3000), "transportOperatorDT"),
    ("pin_lattice_7g", dict(pop=3000), "transportOperatorST"),
]


@pytest.mark.parametrize("name,kw,top", CASES, ids=[f"{c[0]}-{c[2][-2:]}" for c in CASES])
def test_one_cycle_bit_exact(name, kw, top):
    print(f"case {name} {top}", flush=True)
    inp, mg, comp = problem(name, transport=top, **kw)
    pop = kw["pop"]
    master = 0x5C3A84C9
    bank = O.fission_source(comp, pop, master)
    eng = Engine(comp, pop)
    gbank = eng.fission_source(pop, master)
    assert_banks_identical(bank, gbank)
    k = 1.0137
    o_out, o_bins, o_ks, _ = O.eigen_cycle(comp, bank, master + 17, k)
    g_out, g_bins, g_ks = eng.eigen_cycle(bank, master + 17, k)
    assert_banks_identical(o_out, g_out)
    assert np.array_equal(o_ks.view(np.uint64), g_ks.view(np.uint64)), (o_ks, g_ks)
    ok, worst = rel_close(o_bins, g_bins, 1e-12)
    assert ok, f"bins differ, worst rel {worst}"


@pytest.mark.parametrize("excess", [-2500, -700, 0, 900])
def test_norm_size_repr_bit_exact(excess):
    inp, mg, comp = problem("inf_medium_2g", pop=2000)
    eng = Engine(comp, 4000)
    bank = O.fission_source(comp, 2000, 99)
    out, _, _, _ = O.eigen_cycle(comp, bank, 1234, 1.0)
    tot = len(out) - excess
    a = O.norm_size_repr(out, tot, 777)
    b = eng.norm_size_repr(out, tot, 777)
    assert len(a) == tot
    assert_banks_identical(a, b)


@pytest.mark.parametrize("top", ["transportOperatorST", "transportOperatorDT"])
def test_eigen_run_bit_exact(top):
    inp, mg, comp = problem("pin_lattice_7g", pop=2000, transport=top)
    S = dict(pop=2000, inactive=3, active=4, seed=11)
    bank = O.fission_source(comp, 2000, O.rng_skip(11, 0))
    R1, kh1, ba1, _ = O.eigen_run(comp, S, bank)
    eng = Engine(comp, 2000)
    R2, kh2, ba2, _ = eng.eigen_run(S, bank)
    assert np.array_equal(kh1.view(np.uint64), kh2.view(np.uint64)), (kh1, kh2)
    assert list(R1.k_analog) == list(R2.k_analog)
    assert list(R1.k_implicit) == list(R2.k_implicit)
    ok, worst = rel_close(ba1[:, 0], ba2[:, 0], 1e-12)
    assert ok, worst


def test_neutron_mg_imp_cycle_bit_exact():
    """neutronMGimp: the sites keep the colliding particle's collision count, on both sides."""
    print("case neutronMGimp cycle", flush=True)
    from scone_amd import configs, tables as T
    pop = 3000
    inp, mg = configs.pin_lattice_7g(pop=pop, transport="transportOperatorDT")
    inp["collisionOperator"] = {"neutronMG": {"type": "neutronMGimp"}}
    comp = T.compile_problem(inp, mg)
    bank = O.fission_source(comp, pop, 4242)
    eng = Engine(comp, pop)
    o_out, o_bins, o_ks, _ = O.eigen_cycle(comp, bank, 4243, 1.01)
    g_out, g_bins, g_ks = eng.eigen_cycle(bank, 4243, 1.01)
    assert_banks_identical(o_out, g_out)
    assert o_out["collisionN"].max() > 0
    o2, _, _, _ = O.eigen_cycle(comp, o_out, 4244, 1.01)
    g2, _, _ = eng.eigen_cycle(g_out, 4244, 1.01)
    assert_banks_identical(o2, g2)
    eng.close()
