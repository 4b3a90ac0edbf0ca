"""Boundary and tracking branches against the oracle on the GPU, one eigenvalue cycle each, bank bit
for bit and tallies to 1e-12:
  - periodic faces (box_class.f90:380-487 explicitBC / transformBC): the 17x17 lattice with periodic
    x and y faces, surface tracking (explicitBC at each crossing) and delta tracking (transformBC);
  - transportOperatorST with `cache 0` (move_noCache, geometryStd_class.f90:214-276);
  - a zSquareCylinder root border (squareCylinder_class.f90), reflective in x and y, open in z."""
import numpy as np
import pytest

import binding as O
from helpers import assert_banks_identical, rel_close
from scone_amd import configs, tables as T
from scone_amd.engine import Engine

pytestmark = pytest.mark.gpu

POP = 3000


def _cycle(inp, mg, bottom=None, top=None, k=0.97):
    comp = T.compile_problem(inp, mg)
    master = 0x2B7E1516
    bank = O.fission_source(comp, POP, master, bottom, top)
    eng = Engine(comp, POP)
    gbank = eng.fission_source(POP, master, bottom, top)
    assert_banks_identical(bank, gbank)
    o_out, o_bins, o_ks, _ = O.eigen_cycle(comp, bank, master + 5, k)
    g_out, g_bins, g_ks = eng.eigen_cycle(bank, master + 5, k)
    assert_banks_identical(o_out, g_out)
    assert np.array_equal(o_ks.view(np.uint64), g_ks.view(np.uint64)), (o_ks, g_ks)
    ok, worst = rel_close(o_bins, g_bins, 1e-12)
    assert ok, f"bins differ, worst rel {worst}"
    eng.close()
    return o_out


@pytest.mark.parametrize("top", ["transportOperatorST", "transportOperatorDT", "transportOperatorHT"])
def test_periodic_faces_cycle_bit_exact(top):
    print(f"case periodic {top}", flush=True)
    inp, mg = configs.pin_lattice_7g(pop=POP, transport=top)
    inp["geometry"]["boundary"] = [2, 2, 2, 2, 1, 1]
    out = _cycle(inp, mg)
    assert len(out) > 0


def test_surface_tracking_without_cache_bit_exact():
    print("case ST cache 0", flush=True)
    inp, mg = configs.pin_lattice_7g(pop=POP, transport="transportOperatorST")
    inp["transportOperator"]["cache"] = 0
    comp = T.compile_problem(inp, mg)
    assert comp.problem.st_cache == 0
    _cycle(inp, mg)


@pytest.mark.parametrize("top", ["transportOperatorST", "transportOperatorDT"])
def test_square_cylinder_root_border_bit_exact(top):
    print(f"case zSquareCylinder root {top}", flush=True)
    inp, mg = configs.pin_lattice_7g(pop=POP, transport=top)
    g = inp["geometry"]
    g["surfaces"] = {"sq": {"id": 1, "type": "zSquareCylinder", "origin": [0.0, 0.0, 0.0],
                            "halfwidth": [10.71, 10.71, 0.0]}}
    g["boundary"] = [1, 1, 1, 1, 0, 0]
    _cycle(inp, mg, bottom=(-10.71, -10.71, -1.0), top=(10.71, 10.71, 1.0))
