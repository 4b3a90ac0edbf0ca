"""SCONE output files in asciiMATLAB (CPU): the ES12.5 real edit, block prefixes, result arrays
with the leading (value, sd) dimension in Fortran column-major order, and the collisionClerk /
multiMap / spaceMap / keffImplicitClerk layouts (UserInterface/fileOutput/asciiMATLAB_class.f90,
outputFile_class.f90, Tallies/TallyClerks/collisionClerk_class.f90:262-301)."""
import re

import numpy as np
import pytest

from scone_amd import configs, output as OUT, tables as T


@pytest.mark.parametrize("x,s", [
    (1.0, "1.00000E+00"), (-1.0, "-1.00000E+00"), (0.0, "0.00000E+00"), (123456.7, "1.23457E+05"),
    (1.23456789e-5, "1.23457E-05"), (9.999995, "1.00000E+01"), (0.1, "1.00000E-01"),
    (1e100, "1.00000+100"), (-2.5e-120, "-2.50000-120"), (float("nan"), "NaN"),
])
def test_es12_5(x, s):
    assert OUT.es12_5(x) == s
    assert len(s) <= 12


def parse(text):
    """name -> (values, shape) from the asciiMATLAB text."""
    out = {}
    for line in text.strip().split("\n"):
        m = re.match(r"^(\w+) = (.*);$", line)
        assert m, line
        name, rhs = m.groups()
        if rhs.startswith("reshape(["):
            body, dims = re.match(r"reshape\(\[ (.*)\],([\d,]+)\)", rhs).groups()
            shape = [int(d) for d in dims.split(",")]
        elif rhs.startswith("["):
            body, shape = rhs[2:-1], None
        else:
            body, shape = rhs, None
        vals = [v.strip("{}'") for v in body.split(",")]
        out[name] = (vals, shape)
    return out


def test_clerk_results_layout():
    inp, mg = configs.pin_lattice_7g(pop=100)
    comp = T.compile_problem(inp, mg)
    nb = int(comp.problem.n_bins)
    res = np.stack([np.arange(nb) + 0.5, np.arange(nb) * 1e-3], axis=1)
    out = OUT.OutputFile()
    OUT.print_tally(out, comp, res)
    d = parse(out.text())
    assert d["batchSize"] == (["1"], None)
    assert d["pins_g1_multiMapOrder"][0] == ["X", "Y"]
    xb, shape = d["pins_g1_XBounds"]
    assert shape == [2, 17] and float(xb[0]) == -10.71 and float(xb[-1]) == 10.71
    vals, shape = d["pins_g3_Res"]
    assert shape == [2, 2, 17, 17]             # (val, sd) x responses (flux, fiss) x 17 x 17
    a = comp.clerk_info[2]["addr"]
    v = np.array([float(x) for x in vals]).reshape(shape[::-1]).T   # Fortran column-major
    for ix, iy, ir in [(0, 0, 0), (3, 5, 1), (16, 16, 1)]:
        b = a + 2 * (ix + 17 * iy) + ir        # clerk address: width * (bin - 1) + response
        assert v[0, ir, ix, iy] == float(OUT.es12_5(res[b, 0]))
        assert v[1, ir, ix, iy] == float(OUT.es12_5(res[b, 1]))


def test_eigen_results_entries():
    inp, mg = configs.inf_medium_2g(pop=1000, inactive=3, active=7)
    comp = T.compile_problem(inp, mg)

    class R:
        k_implicit = (1.0123, 0.0045)
        imp = [(2.0, 0.1), (0.3, 0.01), (1.5, 0.05), (0.2, 0.02)]

    text = OUT.eigen_results(comp, inp, R, np.ones((2, 2)), cpu_time=3.25, transport_time=2.5).text()
    d = parse(text)
    assert d["pop"][0] == ["1000"] and d["Inactive_Cycles"][0] == ["3"] and d["Active_Cycles"][0] == ["7"]
    assert d["Total_Transport_Time"][0] == ["2.50000E+00"]
    assert d["inactive_batchSize"][0] == ["1"]
    assert d["keff_K_EFF"][0] == ["1.01230E+00", "4.50000E-03"]
    assert d["keff_IMP_ABS"][0] == ["1.50000E+00", "5.00000E-02"]
    assert d["keff_SCATTER_PROD"][0] == ["3.00000E-01", "1.00000E-02"]
    assert d["active_flux_g2_Res"][0] == ["1.00000E+00", "1.00000E+00"]
    order = [line.split(" = ")[0] for line in text.strip().split("\n")]
    assert order[:6] == ["seed", "pop", "Inactive_Cycles", "Active_Cycles", "Total_CPU_Time", "Total_Transport_Time"]
    assert order.index("keff_IMP_PROD") < order.index("keff_IMP_ABS") < order.index("keff_SCATTER_PROD") \
        < order.index("keff_ANA_LEAK") < order.index("keff_K_EFF") < order.index("active_batchSize")


def test_json_printer_layout_by_hand():
    """asciiJSON: entries end with ',', closing a block or the file cuts the last separator,
    arrays open a bracket per dimension that starts at the element (asciiJSON_class.f90)."""
    out = OUT.OutputFile("asciiJSON")
    out.print_value(7, "seed")
    out.start_block("b")
    out.start_array("a", [3])
    for v in (1, 2, 3):
        out.add_value(v)
    out.end_array()
    out.print_result(1.0, 0.5, "r")
    out.end_block()
    assert out.text() == ('{\n  "seed":7,\n  "b":{\n    "a":[1,2,3],\n'
                          '    "r":[1.00000E+00,5.00000E-01]\n  }\n}\n')


def test_json_clerk_results_nest_fastest_dimension_innermost():
    import json
    inp, mg = configs.pin_lattice_7g(pop=100)
    comp = T.compile_problem(inp, mg)
    nb = int(comp.problem.n_bins)
    res = np.stack([np.arange(nb) + 0.5, np.arange(nb) * 1e-3], axis=1)
    out = OUT.OutputFile("asciiJSON")
    OUT.print_tally(out, comp, res)
    d = json.loads(out.text())
    assert d["batchSize"] == 1
    assert d["pins_g1"]["multiMapOrder"] == ["X", "Y"]
    xb = np.array(d["pins_g1"]["XBounds"])
    assert xb.shape == (17, 2) and xb[0, 0] == -10.71 and xb[-1, 1] == 10.71
    v = np.array(d["pins_g3"]["Res"])
    assert v.shape == (17, 17, 2, 2)           # Fortran shape [2, 2, 17, 17], outermost last
    a = comp.clerk_info[2]["addr"]
    for ix, iy, ir in [(0, 0, 0), (3, 5, 1), (16, 16, 1), (9, 2, 0)]:
        b = a + 2 * (ix + 17 * iy) + ir
        assert v[iy, ix, ir, 0] == float(OUT.es12_5(res[b, 0]))
        assert v[iy, ix, ir, 1] == float(OUT.es12_5(res[b, 1]))


def test_json_eigen_results_parse():
    import json
    inp, mg = configs.inf_medium_2g(pop=1000, inactive=3, active=7)
    inp["outputFormat"] = "asciiJSON"
    comp = T.compile_problem(inp, mg)

    class R:
        k_implicit = (1.0123, 0.0045)
        imp = [(2.0, 0.1), (0.3, 0.01), (1.5, 0.05), (0.2, 0.02)]

    out = OUT.eigen_results(comp, inp, R, np.ones((2, 2)), cpu_time=3.25, transport_time=2.5)
    d = json.loads(out.text())
    assert list(d)[:6] == ["seed", "pop", "Inactive_Cycles", "Active_Cycles", "Total_CPU_Time",
                           "Total_Transport_Time"]
    assert d["pop"] == 1000 and d["Total_Transport_Time"] == 2.5
    assert d["inactive"] == {"batchSize": 1}
    assert d["keff"]["K_EFF"] == [1.0123, 0.0045] and d["keff"]["IMP_ABS"] == [1.5, 0.05]
    assert d["active"]["flux_g2"]["Res"] == [[1.0, 1.0]]


def test_output_file_name_and_printer_choice(tmp_path):
    for fmt, ext in (("asciiMATLAB", "m"), ("asciiJSON", "json")):
        out = OUT.OutputFile(fmt)
        out.print_value(1, "x")
        assert out.write(str(tmp_path / "res")) == str(tmp_path / f"res.{ext}")
    out = OUT.OutputFile("dummyPrinter")
    out.start_block("b")
    out.print_result(1.0, 0.1, "r")
    out.end_block()
    assert out.text() == ""
    with pytest.raises(ValueError, match="not recognised"):
        OUT.OutputFile("asciiXML")


def test_output_protocol_errors_and_name_reuse():
    out = OUT.OutputFile()
    out.print_value(1, "x")
    out.print_value(2, "x")
    assert out.warnings and "x is used multiple times" in out.warnings[0]
    with pytest.raises(ValueError):
        out.end_block()
    out.start_array("a", [2])
    out.add_value(1.5)
    with pytest.raises(ValueError, match="mixed"):
        out.add_value("c")
    with pytest.raises(ValueError, match="closed after 1 of 2"):
        out.end_array()
