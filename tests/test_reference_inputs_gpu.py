"""The reference's own benchmark inputs on the MI355X (needs a GPU): C5G7 and SCONE_Slab.

* One cycle of hybrid tracking on the unmodified input: next bank, keffImplicit scores bit-exact
  and the 34x34 fission map to 1e-12 against the oracle.
* The whole power iteration: k_eff against the benchmark's reference eigenvalue for the 2D core,
  1.18655 +/- 0.00008 (OECD/NEA C5G7 MOX benchmark, NEA/NSC/DOC(2003)16, MCNP reference
  solution).  With 2e5 histories per cycle, 40 inactive and 120 active cycles the engine's
  standard deviation is ~1e-4; the check is |k - 1.18655| < 3 sqrt(sd^2 + 0.00008^2) + 0.0002, the
  last term covering cycle-to-cycle correlation that the batch sd does not see.
* SCONE_Slab (URR-2-1-SL, P1 scattering): one cycle bit-exact against the oracle; k within
  0.003 + 3 sd of the analytic 1 (the deck's P1 coefficients exceed 1, where the reference's
  sampler departs from the exact linear kernel: 12 seeds give 1.00101 +/- 0.00011), and delta and
  surface tracking agree within 4 combined sd.
"""
import math
import os

import numpy as np
import pytest

import binding as O
from helpers import assert_banks_identical, rel_close
from scone_amd import dictio, tables as T
from scone_amd.engine import Engine

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
C5G7 = os.path.join(HERE, "golden", "c5g7")
SLAB = os.path.join(HERE, "golden", "scone_slab")
K_REF, K_REF_SD = 1.18655, 0.00008


def c5g7():
    inp = dictio.read(os.path.join(C5G7, "C5G7"))
    return inp, T.compile_problem(inp, base_dir=C5G7)


def test_c5g7_hybrid_cycle_bit_exact():
    print("case c5g7 cycle", flush=True)
    inp, comp = c5g7()
    pop = 20000
    bank = O.fission_source(comp, pop, 3)
    eng = Engine(comp, pop)
    g_bank = eng.fission_source(pop, 3)
    assert_banks_identical(bank, g_bank)
    o_out, o_bins, o_ks, _ = O.eigen_cycle(comp, bank, 4242, 1.18, threads=8)
    g_out, g_bins, g_ks = eng.eigen_cycle(bank, 4242, 1.18)
    assert_banks_identical(o_out, g_out)
    assert np.array_equal(o_ks.view(np.uint64), g_ks.view(np.uint64))
    ok, worst = rel_close(o_bins, g_bins, 1e-12)
    assert ok, worst
    eng.close()


def test_c5g7_keff_matches_benchmark_reference():
    print("case c5g7 keff", flush=True)
    inp, comp = c5g7()
    pop, ni, na = 200000, 40, 120
    eng = Engine(comp, pop)
    bank = eng.fission_source(pop, 11)
    R, kh, bins_a, _ = eng.eigen_run(dict(pop=pop, inactive=ni, active=na, seed=11), bank)
    eng.close()
    k, sd = R.k_implicit
    print(f"k_implicit {k:.6f} +/- {sd:.6f}, k_analog {R.k_analog[0]:.6f}", flush=True)
    tol = 3.0 * math.sqrt(sd * sd + K_REF_SD * K_REF_SD) + 0.0002
    assert abs(k - K_REF) < tol, (k, sd)
    assert bins_a[:, 0].max() > 0


def slab():
    inp = dictio.read(os.path.join(SLAB, "SCONE_Slab"))
    return inp, T.compile_problem(inp, base_dir=SLAB)


def test_scone_slab_p1_cycle_bit_exact():
    print("case slab cycle", flush=True)
    inp, comp = slab()
    pop = 20000
    bank = O.fission_source(comp, pop, 5)
    eng = Engine(comp, pop)
    o_out, o_bins, o_ks, _ = O.eigen_cycle(comp, bank, 777, 1.0, threads=8)
    g_out, g_bins, g_ks = eng.eigen_cycle(bank, 777, 1.0)
    assert_banks_identical(o_out, g_out)
    assert np.array_equal(o_ks.view(np.uint64), g_ks.view(np.uint64))
    ok, worst = rel_close(o_bins, g_bins, 1e-12)
    assert ok, worst
    eng.close()


def test_scone_slab_is_critical_on_the_gpu():
    print("case slab keff", flush=True)
    res = {}
    for top in ("transportOperatorDT", "transportOperatorST"):
        inp = dictio.read(os.path.join(SLAB, "SCONE_Slab"))
        inp["transportOperator"] = {"type": top}
        comp = T.compile_problem(inp, base_dir=SLAB)
        pop, ni, na = 100000, 60, 150
        eng = Engine(comp, pop)
        bank = eng.fission_source(pop, 17)
        R, kh, _, _ = eng.eigen_run(dict(pop=pop, inactive=ni, active=na, seed=17), bank)
        eng.close()
        res[top] = R.k_implicit
        print(f"slab {top} k_implicit {R.k_implicit[0]:.6f} +/- {R.k_implicit[1]:.6f}", flush=True)
    for k, sd in res.values():
        assert abs(k - 1.0) < 0.003 + 3 * sd, (k, sd)
    (k1, s1), (k2, s2) = res.values()
    assert abs(k1 - k2) < 4 * math.sqrt(s1 * s1 + s2 * s2) + 2e-4


def test_package_runs_an_input_file_and_writes_scone_output(tmp_path):
    """scone_amd.package.run_input on the SCONE_Slab input (fewer histories): the result file has
    SCONE's entries; the fission clerk is normalised to normVal 100 every cycle."""
    print("case package run", flush=True)
    from scone_amd import package
    inp = dictio.read(os.path.join(SLAB, "SCONE_Slab"))
    inp["pop"], inp["active"], inp["inactive"], inp["seed"] = 20000, 40, 20, 21
    inp["nuclearData"]["materials"]["fuel"]["xsFile"] = os.path.join(SLAB, "XS", "URRa_2_1_XSS")
    path = tmp_path / "slab_input"
    path.write_text(dictio.dumps(inp))
    out = package.run_input(str(path), str(tmp_path / "slab_out"))
    text = open(out).read()
    lines = dict(l.split(" = ", 1) for l in text.strip().split("\n"))
    assert lines["pop"] == "20000;" and lines["Active_Cycles"] == "40;"
    k, sd = [float(x) for x in lines["keff_K_EFF"].strip("[ ];").split(",")]
    assert abs(k - 1.0) < 4 * sd + 1e-3, (k, sd)
    # a clerk without a map prints Res of shape [2, 1]: the one response, normalised to normVal
    assert lines["active_fiss_Res"] == "reshape([ 1.00000E+02,0.00000E+00],2,1);"


def test_package_writes_ascii_json(tmp_path):
    """The same runner with outputFormat asciiJSON: the file parses as JSON and holds the keff
    block and the clerk results nested as asciiJSON prints them."""
    print("case package run json", flush=True)
    import json
    from scone_amd import package
    inp = dictio.read(os.path.join(SLAB, "SCONE_Slab"))
    inp["pop"], inp["active"], inp["inactive"], inp["seed"] = 5000, 10, 5, 23
    inp["outputFormat"] = "asciiJSON"
    inp["nuclearData"]["materials"]["fuel"]["xsFile"] = os.path.join(SLAB, "XS", "URRa_2_1_XSS")
    path = tmp_path / "slab_input"
    path.write_text(dictio.dumps(inp))
    out = package.run_input(str(path), str(tmp_path / "slab_out"))
    assert out.endswith(".json")
    d = json.load(open(out))
    assert d["pop"] == 5000 and d["Active_Cycles"] == 10
    k, sd = d["keff"]["K_EFF"]
    assert abs(k - 1.0) < 5 * sd + 3e-3, (k, sd)
    assert d["active"]["fiss"]["Res"] == [[100.0, 0.0]]


def test_core_3d_cycle_bit_exact():
    """70-group 3D core (four nesting levels, vacuum, 2e5-bin mesh): one cycle bit-exact."""
    print("case core3d cycle", flush=True)
    from scone_amd import configs
    inp, mg = configs.core_3d(pop=20000, mesh=(40, 40, 5))
    comp = T.compile_problem(inp, mg)
    bank = O.fission_source(comp, 20000, 9)
    eng = Engine(comp, 20000)
    o_out, o_bins, o_ks, _ = O.eigen_cycle(comp, bank, 31337, 0.81, threads=8)
    g_out, g_bins, g_ks = eng.eigen_cycle(bank, 31337, 0.81)
    assert_banks_identical(o_out, g_out)
    assert np.array_equal(o_ks.view(np.uint64), g_ks.view(np.uint64))
    ok, worst = rel_close(o_bins, g_bins, 1e-12)
    assert ok, worst
    eng.close()
