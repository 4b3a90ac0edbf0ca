"""Result files in SCONE's output formats (UserInterface/fileOutput).

`OutputFile` restates outputFile_class.f90: block and array bookkeeping, (value, sd) results whose
array gains a leading dimension 2, reals through the ES12.5 edit and integers through I12, both
left-adjusted, and the name-reuse warning.  It drives one of the printers of asciiOutputFactory
(asciiOutputFactory_func.f90): `asciiMATLAB` (asciiMATLAB_class.f90: block prefixes
`block_name = ...;`, rank > 1 arrays as `reshape([ ... ],d1,d2,...)` in Fortran column-major
order), `asciiJSON` (asciiJSON_class.f90: nested objects, arrays nested from the fastest
dimension inward) and `dummyPrinter` (writes nothing).

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


class _Tape:
    """Character buffer with the delayedStream operations the printers use
    (delayedStream_class.f90:80-111): append, cut the last n characters, peek the n-th last."""

    def __init__(self):
        self.parts = []

    def append(self, s):
        self.parts.append(s)

    def cut(self, n):
        while n > 0 and self.parts:
            last = self.parts.pop()
            if len(last) > n:
                self.parts.append(last[:-n])
                return
            n -= len(last)

    def peek(self, n):
        for part in reversed(self.parts):
            if len(part) >= n:
                return part[-n]
            n -= len(part)
        raise ValueError("peek beyond the start of the buffer")

    def text(self):
        return "".join(self.parts)


class MatlabPrinter:
    """asciiMATLAB printer (asciiMATLAB_class.f90)."""
    ext = "m"
    _BLOCK, _ENTRY, _ARRAY = 0, 1, 2

    def __init__(self):
        self.tape = _Tape()
        self.state = self._BLOCK
        self.prefix = []
        self.shape = []

    def start_block(self, name):
        self.prefix.append(name)

    def end_block(self):
        self.prefix.pop()

    def start_entry(self, name):
        self.state = self._ENTRY
        self.tape.append("".join(p + "_" for p in self.prefix) + name + " = ")

    def end_entry(self):
        self.state = self._BLOCK
        self.tape.append(";\n")

    def start_array(self, shape):
        self.state = self._ARRAY
        self.shape = list(shape)
        self.tape.append("[ " if len(shape) == 1 else "reshape([ ")

    def end_array(self):
        self.state = self._ENTRY
        self.tape.cut(1)                          # the separator after the last element
        self.tape.append("]")
        if len(self.shape) > 1:
            self.tape.append("".join("," + str(d) for d in self.shape) + ")")

    def print_num(self, s):
        if self.state == self._ARRAY:
            self.tape.append(s + ",")
        elif self.state == self._ENTRY:
            self.tape.append(s)

    def print_char(self, s):
        if self.state == self._ARRAY:
            self.tape.append("{'" + s + "'},")
        elif self.state == self._ENTRY:
            self.tape.append("{'" + s + "'}")

    def final_text(self):
        return self.tape.text()


class JsonPrinter:
    """asciiJSON printer (asciiJSON_class.f90): two-space indentation per block level; every
    value is followed by ',', and closing a block or the file cuts the last ',' + newline."""
    ext = "json"
    IND = 2

    def __init__(self):
        self.tape = _Tape()
        self.tape.append("{\n")
        self.level = 1
        self.shape = []
        self.in_array = False
        self.count = 0

    def _cut_separator(self):
        if len(self.tape.parts) and self.tape.peek(2) == ",":
            self.tape.cut(2)
            self.tape.append("\n")

    def start_block(self, name):
        self.tape.append(" " * (self.level * self.IND) + '"' + name + '":{\n')
        self.level += 1

    def end_block(self):
        self._cut_separator()
        self.level -= 1
        self.tape.append(" " * (self.level * self.IND) + "},\n")

    def start_entry(self, name):
        self.tape.append(" " * (self.level * self.IND) + '"' + name + '":')

    def end_entry(self):
        self.tape.append("\n")

    def start_array(self, shape):
        self.shape = list(shape)
        self.in_array = True
        self.count = 0

    def end_array(self):
        self.in_array = False

    def _element(self, s):
        if self.in_array:
            mul = 1
            for d in self.shape:                  # open a bracket per dimension starting here
                mul *= d
                if self.count % mul == 0:
                    self.tape.append("[")
        self.tape.append(s)
        if self.in_array:
            self.count += 1
            mul = 1
            for d in self.shape:                  # close a bracket per dimension ending here
                mul *= d
                if self.count % mul == 0:
                    self.tape.append("]")
        self.tape.append(",")

    def print_num(self, s):
        self._element(s)

    def print_char(self, s):
        self._element('"' + s + '"')

    def final_text(self):
        t = self.tape.text()
        if len(t) >= 2 and t[-2] == ",":
            t = t[:-2] + "\n"
        return t + "}\n"


class DummyPrinter:
    """dummyPrinter (dummyPrinter_class.f90): accepts every call and writes nothing."""
    ext = ""

    def __getattr__(self, name):
        return lambda *a, **k: None

    def final_text(self):
        return ""


PRINTERS = {"asciiMATLAB": MatlabPrinter, "asciiJSON": JsonPrinter, "dummyPrinter": DummyPrinter}


class OutputFile:
    """outputFile (outputFile_class.f90): checks the block/array protocol, converts numbers and
    drives the printer named by `fmt`."""

    _NOT, _UNDEF, _RES, _REAL, _INT, _CHAR = range(6)

    def __init__(self, fmt: str = "asciiMATLAB"):
        if fmt not in PRINTERS:
            raise ValueError(f"output format {fmt} is not recognised; valid: {', '.join(PRINTERS)}")
        self.fmt = fmt
        self.printer = PRINTERS[fmt]()
        self.array = self._NOT
        self.shape = []
        self.top = 0
        self.limit = 0
        self.block = [""]
        self.used = [set()]
        self.warnings = []

    # -- helpers
    def _reuse(self, name):
        if name in self.used[-1]:
            msg = (f"WARNING: Name {name} is used multiple times in block {self.block[-1]} "
                   f"|| Output may not parse correctly!")
            self.warnings.append(msg)
            print(msg)
        self.used[-1].add(name)

    @staticmethod
    def _num(v):
        if isinstance(v, (bool, np.bool_)):
            raise TypeError("logical values are not printed")
        if isinstance(v, (int, np.integer)):
            return str(int(v))                    # I12, adjustl
        return es12_5(float(v))

    # -- outputFile interface
    def start_block(self, name):
        self._reuse(name)
        if self.array != self._NOT:
            raise ValueError(f"cannot start block {name} while writing an array")
        self.used.append(set())
        self.block.append(name)
        self.printer.start_block(name)

    def end_block(self):
        if self.array != self._NOT:
            raise ValueError("cannot exit a block while writing an array")
        if len(self.block) == 1:
            raise ValueError("trying to exit from the root block")
        self.block.pop()
        self.used.pop()
        self.printer.end_block()

    def start_array(self, name, shape):
        self._reuse(name)
        if self.array != self._NOT or len(shape) == 0 or int(np.prod(shape)) == 0:
            raise ValueError(f"cannot start array {name}")
        self.shape = [int(d) for d in shape]
        self.array = self._UNDEF
        self.printer.start_entry(name)

    def end_array(self):
        if self.top != self.limit:
            raise ValueError(f"array closed after {self.top} of {self.limit} entries")
        self.array, self.top, self.limit = self._NOT, 0, 0
        self.printer.end_array()
        self.printer.end_entry()

    def add_result(self, val, std):
        if self.array == self._UNDEF:
            self.array, self.top = self._RES, 0
            self.limit = 2 * int(np.prod(self.shape))
            self.printer.start_array([2] + self.shape)
        if self.array != self._RES or self.top + 2 > self.limit:
            raise ValueError("result does not fit the array")
        self.printer.print_num(es12_5(float(val)))
        self.printer.print_num(es12_5(float(std)))
        self.top += 2

    def add_value(self, v):
        kind = self._CHAR if isinstance(v, str) else \
            self._INT if isinstance(v, (int, np.integer)) else self._REAL
        if self.array == self._UNDEF:
            self.array, self.top = kind, 0
            self.limit = int(np.prod(self.shape))
            self.printer.start_array(self.shape)
        if self.array != kind or self.top + 1 > self.limit:
            raise ValueError("value does not fit the array (mixed content or overflow)")
        if kind == self._CHAR:
            self.printer.print_char(v)
        else:
            self.printer.print_num(self._num(v))
        self.top += 1

    def print_value(self, v, name):
        self._reuse(name)
        self.printer.start_entry(name)
        if isinstance(v, str):
            self.printer.print_char(v)
        else:
            self.printer.print_num(self._num(v))
        self.printer.end_entry()

    def print_result(self, val, std, name):
        self._reuse(name)
        self.printer.start_entry(name)
        self.printer.start_array([2])
        self.printer.print_num(es12_5(float(val)))
        self.printer.print_num(es12_5(float(std)))
        self.printer.end_array()
        self.printer.end_entry()

    def text(self) -> str:
        """The file's contents as closing it would leave them (the printer's endFile)."""
        return self.printer.final_text()

    def write(self, path):
        """Write to `path` + '.' + the printer's extension, as outputFile%init names the file."""
        path = f"{path}.{self.printer.ext}"
        with open(path, "w") as f:
            f.write(self.text())
        return path


def _print_map(out, axes, multi):
    """multiMap%print / spaceMap%print / materialMap%print / energyMap%print / cellMap%print."""
    names = []
    for a in axes:
        names.append({"space": a[1].upper() if a[0] == "space" else "", "material": "Material",
                      "energy": "Energy", "cell": "Cell"}[a[0]])
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
        elif a[0] == "energy":                     # lower bounds, then upper bounds
            edges = a[1]
            N = len(edges) - 1
            out.start_array(n + "Bounds", [N, 2])
            for i in range(N):
                out.add_value(float(edges[i]))
            for i in range(N):
                out.add_value(float(edges[i + 1]))
            out.end_array()
        elif a[0] == "cell":                       # cell IDs as text, then 'undefined'
            clist, undef = a[1], a[2]
            out.start_array(n + "Bins", [1, len(clist) + (1 if undef else 0)])
            for c in clist:
                out.add_value(str(int(c)))
            if undef:
                out.add_value("undefined")
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

