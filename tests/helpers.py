"""Shared fixtures for the parity tests: problems, banks and comparisons."""
import numpy as np

from scone_amd import configs, tables as T

BANK_FIELDS = ["wgt", "r", "dir", "E", "G", "isMG", "type", "time", "lambda", "matIdx", "cellIdx",
               "uniqueID", "collisionN", "broodID"]


def problem(name, **kw):
    inp, mg = configs.CONFIGS[name](**kw)
    return inp, mg, T.compile_problem(inp, mg)


def assert_banks_identical(a, b, fields=None):
    assert len(a) == len(b), f"bank sizes differ: {len(a)} vs {len(b)}"
    for f in fields or BANK_FIELDS:
        x, y = a[f], b[f]
        if x.dtype.kind == "f":
            bad = np.nonzero(x.view(np.uint64) != y.view(np.uint64))[0] if x.ndim == 1 else \
                np.nonzero((x.view(np.uint64) != y.view(np.uint64)).any(axis=1))[0]
        else:
            bad = np.nonzero(x != y)[0] if x.ndim == 1 else np.nonzero((x != y).any(axis=1))[0]
        assert len(bad) == 0, f"field {f} differs at {len(bad)} sites, first {bad[:5]}: {x[bad[:3]]} vs {y[bad[:3]]}"


def rel_close(a, b, rtol=1e-12):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    scale = np.maximum(np.abs(a), np.abs(b))
    diff = np.abs(a - b)
    ok = diff <= rtol * scale
    return bool(ok.all()), float(np.max(np.where(scale > 0, diff / np.where(scale > 0, scale, 1), 0.0)))


def dummy_mg(nG=1):
    return {"numberOfGroups": nG, "capture": [0.5] * nG, "scatteringMultiplicity": [1.0] * (nG * nG),
            "P0": [0.5 / nG] * (nG * nG)}


def geometry_problem(universes, surfaces, boundary=(0, 0, 0, 0, 0, 0), materials=("m1",), graph="shrunk",
                     cells=None):
    """Compile a geometry-only SCONE input (1-group dummy materials) for geometry unit tests."""
    inp = {
        "transportOperator": {"type": "transportOperatorST"},
        "geometry": {"type": "geometryStd", "boundary": list(boundary), "graph": {"type": graph},
                     "surfaces": surfaces, "cells": cells or {}, "universes": universes},
        "nuclearData": {"handles": {"mg": {"type": "baseMgNeutronDatabase", "PN": "P0"}},
                        "materials": {m: {"xsFile": m} for m in materials}},
        "activeTally": {},
    }
    return T.compile_problem(inp, {m: dummy_mg() for m in materials})


def assert_results_close(ref, got, n_batches, rtol=1e-12, label=""):
    """(mean, sd) score results that differ only by summation order: means within rtol; each sd
    within rtol times the condition number of scoreMemory%getResult's
    sd = sqrt(CSUM2/(N(N-1)) - mean^2/(N-1)), which cancels:
    cond = (1 + 3 mean^2 / ((N-1) sd^2)) / 2."""
    ref = np.asarray(ref, dtype=float).reshape(-1, 2)
    got = np.asarray(got, dtype=float).reshape(-1, 2)
    ok, worst = rel_close(ref[:, 0], got[:, 0], rtol)
    assert ok, f"{label}: means differ, worst rel {worst}"
    N = n_batches
    if N < 2:
        return
    m, s = ref[:, 0], ref[:, 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        cond = np.where(s > 0, 0.5 * (1.0 + 3.0 * m * m / ((N - 1) * s * s)), 1.0)
    tol = rtol * np.maximum(cond, 1.0) * np.maximum(np.abs(s), np.abs(got[:, 1]))
    bad = np.nonzero(np.abs(s - got[:, 1]) > tol + 1e-300)[0]
    assert len(bad) == 0, f"{label}: sd differs beyond its conditioning at bins {bad[:5]}: {s[bad[:3]]} vs {got[bad[:3], 1]}"
