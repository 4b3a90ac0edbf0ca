"""Parity of the HIP engines with the oracle on CSG-cell geometries (needs a GPU).

Cell universes of every surface type (truncated cone and wedge included), a rotated universe, a 'global' universe at five nesting
levels, a sphere as the root border and track-length clerks run through the general kernels; the
contract is the one of the lattice/pin geometries: per-history results bit-exact, clerk scores
within 1e-12 relative (summation order)."""
import numpy as np
import pytest

import binding as O
from helpers import assert_banks_identical, assert_results_close, rel_close
from scone_amd import configs, tables as T
from scone_amd.engine import Batch, Engine, secondary_bytes

pytestmark = pytest.mark.gpu

TOPS = ["transportOperatorST", "transportOperatorDT", "transportOperatorHT"]


@pytest.mark.parametrize("variant", ["3-level", "5-level", "cone-wedge-union"])
@pytest.mark.parametrize("top", TOPS, ids=["ST", "DT", "HT"])
def test_cell_lattice_cycle_bit_exact(top, variant):
    print(f"case cell_lattice {top} {variant}", flush=True)
    pop = 3000
    inp, mg = configs.cell_lattice(pop=pop, transport=top, outer=variant == "5-level",
                                   shapes=variant == "cone-wedge-union")
    comp = T.compile_problem(inp, mg)
    master = 0x5C3A84C9
    bank = O.fission_source(comp, pop, master)
    eng = Engine(comp, pop)
    gbank = eng.fission_source(pop, master)
    assert_banks_identical(bank, gbank)
    k = 0.93
    o_out, o_bins, o_ks, _ = O.eigen_cycle(comp, bank, master + 17, k)
    g_out, g_bins, g_ks = eng.eigen_cycle(bank, master + 17, k)
    assert_banks_identical(o_out, g_out)
    assert np.array_equal(o_ks.view(np.uint64), g_ks.view(np.uint64))
    ok, worst = rel_close(o_bins, g_bins, 1e-12)
    assert ok, f"bins differ, worst rel {worst}"
    eng.close()


def test_cell_lattice_eigen_run_bit_exact():
    print("case cell_lattice eigen run DT", flush=True)
    inp, mg = configs.cell_lattice(pop=2000, transport="transportOperatorDT")
    comp = T.compile_problem(inp, mg)
    S = dict(pop=2000, inactive=3, active=4, seed=11)
    bank = O.fission_source(comp, 2000, O.rng_skip(11, 0))
    R1, kh1, ba1, _ = O.eigen_run(comp, S, bank)
    eng = Engine(comp, 2000)
    R2, kh2, ba2, _ = eng.eigen_run(S, bank)
    assert np.array_equal(kh1.view(np.uint64), kh2.view(np.uint64)), (kh1, kh2)
    assert list(R1.k_implicit) == list(R2.k_implicit)
    ok, worst = rel_close(ba1[:, 0], ba2[:, 0], 1e-12)
    assert ok, worst
    eng.close()


def _fixed_check(B, slot, comp, src, pop, cycles, seed, buffer, label):
    res, done, h = O.fixed_run(comp, src, pop, cycles, seed, hist=True, threads=8, buffer=buffer)
    r, bins, hr, hc = B.query(slot)
    assert r.state == T.SLOT_DONE and r.n_batches == done, label
    bad = np.nonzero(hr != h["rng_end"])[0]
    assert len(bad) == 0, f"{label}: {len(bad)} histories end on another RNG state, first {bad[:5]}"
    assert np.array_equal(hc, h["n_coll"]), label
    assert_results_close(res, bins, done, 1e-12, label)


@pytest.mark.parametrize("top", TOPS[:2], ids=["ST", "DT"])
def test_fixed_source_cells_and_sphere_border_in_one_batch(top):
    """A fissile cell lattice (secondaries in the history) and an absorbing sphere (non-box root
    border) resident together with a lattice/pin problem: the batch runs the general kernels for
    all three and each matches its oracle run."""
    print(f"case fixed cells {top}", flush=True)
    probs = []
    inp, mg = configs.cell_lattice(pop=1200, active=3, transport=top, fixed=True, shapes=True)
    probs.append((inp, mg, 1200, 3, inp["buffer"]))
    inp, mg = configs.absorbing_sphere(pop=1500, cycles=3, transport=top)
    probs.append((inp, mg, 1500, 3, 50))
    inp, mg = configs.slab_fixed(pop=1000, cycles=3, transport=top)
    probs.append((inp, mg, 1000, 3, 50))
    tcode = T.TRANSPORT_ST if top.endswith("ST") else T.TRANSPORT_DT
    B = Batch(transport=tcode, max_slots=4,
              hbm_budget=(512 << 20) + secondary_bytes(tcode, max(p[4] for p in probs)))
    slots = []
    for inp, mg, pop, cycles, buf in probs:
        comp = T.compile_problem(inp, mg)
        src = T.compile_source(inp["source"], comp)
        slots.append((B.add(comp, src, pop, cycles, inp["seed"], record=True, buffer=buf), comp, src, pop, cycles,
                      inp["seed"], buf))
    B.run_all()
    for i, (s, comp, src, pop, cycles, seed, buf) in enumerate(slots):
        _fixed_check(B, s, comp, src, pop, cycles, seed, buf, f"{top} problem {i}")
    B.close()
