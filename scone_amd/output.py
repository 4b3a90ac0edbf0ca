"""Result files in SCONE's output format (UserInterface/fileOutput).

`OutputFile` restates outputFile_class.f90 with the asciiMATLAB printer (asciiMATLAB_class.f90):
blocks prefix entry names (`block_name = ...;`), results are (value, sd) pairs whose array gains
a leading dimension 2, rank > 1 arrays are written `reshape([ ... ],d1,d2,...)` in Fortran
column-major order, reals use the ES12.5 edit descriptor (left-adjusted), integers I12.

`print_tally` restates tallyAdmin%print (Tallies/tallyAdmin_class.f90:436-451) with
collisionClerk%print (collisionClerk_class.f90:262-301) and the map prints of multiMap,
spaceMap and materialMap; `eigen_results` / `fixed_results` restate the packages' collectResults
(eigenPhysicsPackage_class.f90:376-419, fixedSourcePhysicsPackage_class.f90:293-319).
"""
from __future__ import annotations

import math

import numpy as np


def es12_5(x: float) -> str:
    """Fortran ES12.5 edit of a real, left-adjusted: d.ddddd followed by the exponent 'E+dd'
    (three-digit exponents drop the 'E'); NaN and infinities as gfortran writes them."""
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    s = f"{x:.5E}"                      # correctly rounded, like gfortran's formatted output
    mant, exp = s.split("E")
    e = int(exp)
    if abs(e) <= 99:
        return f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}"
    return f"{mant}{'+' if e >= 0 else '-'}{abs(e):03d}"


class OutputFile:
    """outputFile + asciiMATLAB printer."""

    def __init__(self, fmt: str = "asciiMATLAB"):
        if fmt != "asciiMATLAB":
            raise NotImplementedError(f"output format {fmt}: only asciiMATLAB is written")
        self.fmt = fmt
        self.buf = []
        self.prefix = []
        self.shape = None
        self.array = None            # None, "undef", "res", "val"
        self.top = 0
        self.limit = 0

    # -- asciiMATLAB printer
    def _entry(self, name):
        self.buf.append("".join(p + "_" for p in self.prefix) + name + " = ")

    def _start_array(self, shape):
        self.shape = list(shape)
        self.buf.append("[ " if len(shape) == 1 else "reshape([ ")

    def _end_array(self):
        last = self.buf.pop()
        self.buf.append(last[:-1])               # cut the trailing ','
        self.buf.append("]")
        if len(self.shape) > 1:
            self.buf.append("".join("," + str(d) for d in self.shape) + ")")

    def _num(self, s, in_array):
        self.buf.append(s + "," if in_array else s)

    def _char(self, s, in_array):
        self.buf.append("{'" + s + "'}" + ("," if in_array else ""))

    # -- outputFile interface
    def start_block(self, name):
        if self.array is not None:
            raise ValueError(f"cannot start block {name} while writing an array")
        self.prefix.append(name)

    def end_block(self):
        if not self.prefix:
            raise ValueError("trying to exit from the root block")
        self.prefix.pop()

    def start_array(self, name, shape):
        if self.array is not None or len(shape) == 0 or int(np.prod(shape)) == 0:
            raise ValueError(f"cannot start array {name}")
        self.shape = list(shape)
        self.array = "undef"
        self._entry(name)

    def add_result(self, val, std):
        if self.array == "undef":
            self.array = "res"
            self.top = 0
            self.limit = 2 * int(np.prod(self.shape))
            self._start_array([2] + self.shape)
        if self.array != "res" or self.top + 2 > self.limit:
            raise ValueError("result does not fit the array")
        self._num(es12_5(val), True)
        self._num(es12_5(std), True)
        self.top += 2

    def add_value(self, v):
        if self.array == "undef":
            self.array = "val"
            self.top = 0
            self.limit = int(np.prod(self.shape))
            self._start_array(self.shape)
        if self.array != "val" or self.top + 1 > self.limit:
            raise ValueError("value does not fit the array")
        if isinstance(v, str):
            self._char(v, True)
        elif isinstance(v, (int, np.integer)):
            self._num(str(int(v)), True)
        else:
            self._num(es12_5(float(v)), True)
        self.top += 1

    def end_array(self):
        if self.top != self.limit:
            raise ValueError(f"array closed after {self.top} of {self.limit} entries")
        self._end_array()
        self.buf.append(";\n")
        self.array = None

    def print_value(self, v, name):
        self._entry(name)
        if isinstance(v, str):
            self._char(v, False)
        elif isinstance(v, (int, np.integer)):
            self._num(str(int(v)), False)
        else:
            self._num(es12_5(float(v)), False)
        self.buf.append(";\n")

    def print_result(self, val, std, name):
        self._entry(name)
        self._start_array([2])
        self._num(es12_5(val), True)
        self._num(es12_5(std), True)
        self._end_array()
        self.buf.append(";\n")

    def text(self) -> str:
        return "".join(self.buf)

    def write(self, path):
        if not path.endswith(".m"):
            path = path + ".m"
        with open(path, "w") as f:
            f.write(self.text())
        return path


def _print_map(out, axes, multi):
    """multiMap%print / spaceMap%print / materialMap%print."""
    names = []
    for a in axes:
        names.append(a[1].upper() if a[0] == "space" else "Material")
    if multi:
        out.start_array("multiMapOrder", [len(axes)])
        for n in names:
            out.add_value(n)
        out.end_array()
    for a, n in zip(axes, names):
        if a[0] == "space":
            edges = a[2]
            out.start_array(n + "Bounds", [2, len(edges) - 1])
            for i in range(len(edges) - 1):
                out.add_value(float(edges[i]))
                out.add_value(float(edges[i + 1]))
            out.end_array()
        else:
            mlist, undef = a[1], a[2]
            out.start_array("MaterialBins", [1, len(mlist) + (1 if undef else 0)])
            for m in mlist:
                out.add_value(m)
            if undef:
                out.add_value("undefined")
            out.end_array()


def print_tally(out, comp, results, batch_size=1):
    """tallyAdmin%print: batchSize, then every collision clerk's block.  results: (n_bins, 2)
    array of (mean, sd) in score-memory order."""
    out.print_value(int(batch_size), "batchSize")
    res = np.asarray(results, dtype=float).reshape(-1, 2)
    for info in comp.clerk_info:
        out.start_block(info["name"])
        if info["axes"]:
            _print_map(out, info["axes"], info.get("multi", False))
        out.start_array("Res", info["shape"])
        a = info["addr"]
        for i in range(info["size"]):
            out.add_result(res[a + i, 0], res[a + i, 1])
        out.end_array()
        out.end_block()


def eigen_results(comp, inp, R, bins_active, bins_inactive=None, cpu_time=0.0, transport_time=0.0):
    """eigenPhysicsPackage%collectResults with the engine's eigen result R (scn_eigen_result)."""
    out = OutputFile(inp.get("outputFormat", "asciiMATLAB"))
    out.print_value(int(inp.get("seed", 0)), "seed")
    out.print_value(int(inp["pop"]), "pop")
    out.print_value(int(inp["inactive"]), "Inactive_Cycles")
    out.print_value(int(inp["active"]), "Active_Cycles")
    out.print_value(float(cpu_time), "Total_CPU_Time")
    out.print_value(float(transport_time), "Total_Transport_Time")
    out.start_block("inactive")
    # the inactive tally: its clerks (the active ones) when they scored inactive cycles, else none
    if bins_inactive is None:
        out.print_value(1, "batchSize")
    else:
        print_tally(out, comp, bins_inactive)
    out.end_block()
    # activeAtch: the keffImplicitClerk named 'keff' (eigenPhysicsPackage_class.f90:620-633)
    out.print_value(1, "batchSize")
    out.start_block("keff")
    for name, k in (("IMP_PROD", 0), ("IMP_ABS", 2), ("SCATTER_PROD", 1), ("ANA_LEAK", 3)):
        out.print_result(R.imp[k][0], R.imp[k][1], name)
    out.print_result(R.k_implicit[0], R.k_implicit[1], "K_EFF")
    out.end_block()
    out.start_block("active")
    print_tally(out, comp, bins_active)
    out.end_block()
    return out


def fixed_results(comp, inp, bins, seed, pop, n_batches, cpu_time=0.0, transport_time=0.0):
    """fixedSourcePhysicsPackage%collectResults."""
    out = OutputFile(inp.get("outputFormat", "asciiMATLAB"))
    out.print_value(int(seed), "seed")
    out.print_value(int(pop), "pop")
    out.print_value(int(n_batches), "Source_batches")
    out.print_value(float(cpu_time), "Total_CPU_Time")
    out.print_value(float(transport_time), "Transport_time")
    print_tally(out, comp, bins)
    return out

