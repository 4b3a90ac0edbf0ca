"""The reference's own C5G7 benchmark input (InputFiles/Benchmarks/Multigroup/C5G7 with its
XS_C5G7 decks, kept as data fixtures under tests/golden/c5g7) through the dictionary reader and
the table compiler (CPU): hybrid tracking (transportOperatorHT, cutoff 0.9), extended graph,
three nesting levels of lattices and pins, the 34x34 fission-rate map.

Placement is checked at pin centres read off the maps (latCore 10/20/30, latUO2, latMOX, written
as seen from +z), and the HT selection rule against its two limits: in a single-material medium
sigmaT / majorant = 1 > 1 - cutoff (always delta tracking); with cutoff 0 the test ratio > 1 never
holds (always surface tracking).  Each limit must reproduce the plain operator bit for bit.
"""
import os

import numpy as np
import pytest

import binding as O
from helpers import assert_banks_identical, problem
from scone_amd import dictio, tables as T

HERE = os.path.dirname(os.path.abspath(__file__))
C5G7 = os.path.join(HERE, "golden", "c5g7")
MAT = {"mox43": 1, "mox7": 2, "mox87": 3, "UO2": 4, "FC": 5, "GT": 6, "water": 7}


def c5g7():
    inp = dictio.read(os.path.join(C5G7, "C5G7"))
    return inp, T.compile_problem(inp, base_dir=C5G7)


def test_c5g7_compiles_as_the_reference_reads_it():
    inp, comp = c5g7()
    P = comp.problem
    assert P.transport == T.TRANSPORT_HT and P.ht_cutoff == 0.9
    assert comp.mat_names == list(MAT)
    assert P.n_bins == 34 * 34
    assert comp.bounds == pytest.approx([-32.13] * 3 + [32.13] * 3)
    assert list(comp.problem.boxes[P.border_box].bc) == [1, 0, 0, 1, 1, 1]
    assert P.mg.n_groups == 7
    # UO2: nu*sigma_f of group 1 from the deck (nu 2.7814494, fission 7.212060E-3)
    x = comp.xs[MAT["UO2"] - 1, 0]
    assert x[4] == 2.7814494 * 7.212060e-3


@pytest.mark.parametrize("r,mat", [
    ((-31.5, 31.5, 0.0), "UO2"),            # UO2 assembly (core row 1, col 1), pin (1, 1)
    ((-30.9, 31.5, 0.0), "water"),          # same pin cell, outside r = 0.54
    ((-25.2, 28.98, 0.0), "GT"),            # UO2 assembly pin (3, 6): guide tube
    ((-21.42, 21.42, 0.0), "FC"),           # UO2 assembly centre: fission chamber
    ((-10.08, 31.5, 0.0), "mox43"),         # MOX assembly (core row 1, col 2), pin (1, 1)
    ((-8.82, 30.24, 0.0), "mox7"),          # MOX pin (2, 2)
    ((-3.78, 27.72, 0.0), "mox87"),         # MOX pin (4, 6)
    ((-31.5, 10.08, 0.0), "mox43"),         # MOX assembly (core row 2, col 1), pin (1, 1)
    ((0.0, 0.0, 0.0), "FC"),                # core centre: UO2 assembly centre pin
    ((-20.0, -20.0, 0.0), "water"),         # core row 3: water assemblies
    ((20.0, 20.0, 5.0), "water"),           # core col 3: water assemblies
])
def test_c5g7_material_at_pin_centres(r, mat):
    _, comp = c5g7()
    out = O.place(comp, r)
    assert out.matIdx == MAT[mat], (r, out.matIdx)
    # root -> core lattice -> assembly lattice -> pin; the water assemblies are a pin universe
    assert out.nesting == (3 if r[0] > 10.71 or r[1] < -10.71 else 4)


@pytest.mark.parametrize("cutoff,plain", [(0.9, "transportOperatorDT"), (0.0, "transportOperatorST")])
def test_hybrid_tracking_limits_reproduce_plain_operators(cutoff, plain):
    pop = 1500
    _, _, comp_ht = problem("inf_medium_2g", pop=pop, transport="transportOperatorHT")
    comp_ht.problem.ht_cutoff = cutoff
    _, _, comp_p = problem("inf_medium_2g", pop=pop, transport=plain)
    bank = O.fission_source(comp_p, pop, 77)
    a = O.eigen_cycle(comp_ht, bank, 1234, 1.01, hist=True)
    b = O.eigen_cycle(comp_p, bank, 1234, 1.01, hist=True)
    assert_banks_identical(a[0], b[0])
    assert np.array_equal(a[3]["rng_end"], b[3]["rng_end"])
    assert np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))


def test_c5g7_hybrid_cycle_runs_on_the_oracle():
    inp, comp = c5g7()
    pop = 3000
    bank = O.fission_source(comp, pop, 5)
    out, bins, ks, h = O.eigen_cycle(comp, bank, 99, 1.18, hist=True, threads=4)
    k_est = O.keff_implicit(ks)
    assert 0.9 < k_est < 1.5
    assert len(out) > 0.7 * pop and bins.sum() > 0


def test_core_3d_config_nests_four_levels():
    """The 3D core configuration: pin -> assembly -> core lattice -> root, reflector ring as the
    core lattice's padding, vacuum box 400 cm tall."""
    from scone_amd import configs
    inp, mg = configs.core_3d(pop=100, mesh=(10, 10, 5))
    comp = T.compile_problem(inp, mg)
    assert comp.nG == 70 and comp.problem.n_bins == 70 * 10 * 10 * 5
    assert comp.bounds == pytest.approx([-180.65, -180.65, -200.0, 180.65, 180.65, 200.0])
    names = comp.mat_names
    assert O.place(comp, (0.0, 0.0, 0.0)).matIdx == names.index("structural") + 1   # core centre pin: guide tube
    assert O.place(comp, (0.6, 0.0, 0.0)).matIdx == names.index("moderator") + 1
    assert O.place(comp, (0.0 - 1.26, 0.0, 150.0)).matIdx == names.index("fuel") + 1
    out = O.place(comp, (170.0, 0.0, 0.0))
    assert out.matIdx == names.index("reflector") + 1 and out.nesting == 2
