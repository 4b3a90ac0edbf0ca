"""Synthetic multigroup problems in SCONE's input-dictionary form.

Each builder returns ``(input_dict, mg_dicts)``: the eigenPhysicsPackage dictionary (keys as in
InputFiles/SCONE_Inf and InputFiles/Benchmarks/Multigroup/C5G7) and the per-material MG data
dictionaries (the xsFile format of IntegrationTestFiles/mgMat2).  Data come from numpy's PCG64
generator with a fixed seed, so a problem is reproducible from its name alone.

* ``inf_medium_2g``   -- 2-group infinite medium: reflective 10 cm cube, one material,
                         Sigma_t ~ U[0.5,2], c ~ U[0.6,0.9], full 2x2 P0, nu scaled to k_inf = 1,
                         chi = [1, 0], surface tracking, per-group collision-estimator flux.
* ``pin_lattice_7g``  -- 17x17 square lattice of cylindrical pins (pitch 1.26, r 0.54) with the
                         C5G7 UO2 assembly pin map, reflective box, 7 groups, fuel / structural /
                         moderator materials, per-pin x per-group flux and fission-rate clerks.
"""
from __future__ import annotations

import numpy as np

# C5G7 UO2 assembly layout (InputFiles/Benchmarks/Multigroup/C5G7 latUO2): 1 fuel, 2 guide tube,
# 6 fission chamber; the map is written as it looks from +z (top row = +y).
UO2_MAP = [
    "1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1",
    "1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1",
    "1 1 1 1 1 2 1 1 2 1 1 2 1 1 1 1 1",
    "1 1 1 2 1 1 1 1 1 1 1 1 1 2 1 1 1",
    "1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1",
    "1 1 2 1 1 2 1 1 2 1 1 2 1 1 2 1 1",
    "1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1",
    "1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1",
    "1 1 2 1 1 2 1 1 6 1 1 2 1 1 2 1 1",
    "1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1",
    "1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1",
    "1 1 2 1 1 2 1 1 2 1 1 2 1 1 2 1 1",
    "1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1",
    "1 1 1 2 1 1 1 1 1 1 1 1 1 2 1 1 1",
    "1 1 1 1 1 2 1 1 2 1 1 2 1 1 1 1 1",
    "1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1",
    "1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1 1",
]


def _mg_dict(sig_t, P0, capture, fission=None, nu=None, chi=None):
    """MG data dictionary; P0[g_out][g_in] written one line per incident group (mgMat2 format)."""
    nG = len(sig_t)
    d = {"numberOfGroups": nG, "capture": [float(x) for x in capture]}
    d["scatteringMultiplicity"] = [1.0] * (nG * nG)
    d["P0"] = [float(P0[gout][gin]) for gin in range(nG) for gout in range(nG)]
    if fission is not None:
        d["fission"] = [float(x) for x in fission]
        d["nu"] = [float(x) for x in nu]
        d["chi"] = [float(x) for x in chi]
    return d


def k_inf_and_flux(mg: dict):
    """Infinite-medium k_inf = nuSf^T (T - S)^-1 chi and the per-source-neutron group flux,
    from the data dictionary as the material init would assemble it."""
    nG = mg["numberOfGroups"]
    P0 = np.array(mg["P0"]).reshape(nG, nG).T          # P0[g_out, g_in]
    scat = P0.sum(axis=0)
    fis = np.array(mg.get("fission", [0.0] * nG))
    tot = scat + np.array(mg["capture"]) + fis
    chi = np.array(mg.get("chi", [0.0] * nG))
    chi = chi / chi.sum() if chi.sum() > 0 else chi
    nu = np.array(mg.get("nu", [0.0] * nG))
    A = np.diag(tot) - P0
    phi = np.linalg.solve(A, chi)
    return float((nu * fis) @ phi), phi


def inf_medium_2g(seed: int = 1, pop: int = 10000, inactive: int = 20, active: int = 50,
                  transport: str = "transportOperatorST"):
    rng = np.random.default_rng(seed)
    nG = 2
    sig_t = rng.uniform(0.5, 2.0, nG)
    c = rng.uniform(0.6, 0.9, nG)
    raw = rng.uniform(0.0, 1.0, (nG, nG))              # raw[g_out, g_in]
    P0 = np.zeros((nG, nG))
    for g in range(nG):
        P0[:, g] = raw[:, g] / raw[:, g].sum() * c[g] * sig_t[g]
    fis = 0.5 * (1.0 - c) * sig_t
    cap = (1.0 - c) * sig_t - fis
    chi = np.array([1.0, 0.0])
    mg = _mg_dict(sig_t, P0, cap, fis, np.ones(nG), chi)
    k1, _ = k_inf_and_flux(mg)
    mg["nu"] = [1.0 / k1] * nG
    tallies = {}
    for g in range(1, nG + 1):
        tallies[f"flux_g{g}"] = {"type": "collisionClerk", "filter": {"type": "energyFilter", "Gtop": g, "Glow": g},
                                 "response": ["flux"], "flux": {"type": "fluxResponse"}}
    inp = {
        "type": "eigenPhysicsPackage", "pop": pop, "active": active, "inactive": inactive,
        "XSdata": "mg", "dataType": "mg", "seed": seed,
        "collisionOperator": {"neutronMG": {"type": "neutronMGstd"}},
        "transportOperator": {"type": transport},
        "inactiveTally": {}, "activeTally": tallies,
        "geometry": {"type": "geometryStd", "boundary": [1, 1, 1, 1, 1, 1], "graph": {"type": "shrunk"},
                     "surfaces": {"cube": {"id": 1, "type": "box", "origin": [0.0, 0.0, 0.0],
                                           "halfwidth": [5.0, 5.0, 5.0]}},
                     "cells": {},
                     "universes": {"root": {"id": 1, "type": "rootUniverse", "border": 1, "fill": "medium"}}},
        "nuclearData": {"handles": {"mg": {"type": "baseMgNeutronDatabase", "PN": "P0"}},
                        "materials": {"medium": {"temp": 300, "composition": {}, "xsFile": "medium"}}},
    }
    return inp, {"medium": mg}


def _lower_tri_scatter(rng, nG, sig_t, c, self_frac_lo, self_frac_hi):
    """P0[g_out, g_in] = 0 for g_out < g_in; within-group share in [lo, hi], rest downscatter."""
    P0 = np.zeros((nG, nG))
    for g in range(nG):
        s = c[g] * sig_t[g]
        if g == nG - 1:
            P0[g, g] = s
            continue
        selff = rng.uniform(self_frac_lo, self_frac_hi)
        down = rng.uniform(0.0, 1.0, nG - g - 1) * np.exp(-np.arange(nG - g - 1))
        down = down / down.sum()
        P0[g, g] = selff * s
        P0[g + 1:, g] = (1.0 - selff) * s * down
    return P0


def pin_lattice_materials(seed: int = 2, nG: int = 7):
    rng = np.random.default_rng(seed)
    # fuel: fissile, moderate scattering, fission peaking in the thermal groups
    st_f = np.concatenate([rng.uniform(0.2, 0.6, 3), rng.uniform(0.5, 1.5, nG - 3)])
    c_f = np.concatenate([rng.uniform(0.75, 0.95, 3), rng.uniform(0.4, 0.7, nG - 3)])
    P0_f = _lower_tri_scatter(rng, nG, st_f, c_f, 0.5, 0.9)
    absr = (1.0 - c_f) * st_f
    ff = np.concatenate([rng.uniform(0.1, 0.3, 3), rng.uniform(0.4, 0.7, nG - 3)])
    fis_f = ff * absr
    cap_f = absr - fis_f
    chi = np.zeros(nG)
    chi[:3] = rng.uniform(0.0, 1.0, 3)
    chi /= chi.sum()
    nu = np.full(nG, 2.43)
    fuel = _mg_dict(st_f, P0_f, cap_f, fis_f, nu, chi)
    # structural (guide tube / chamber): no fission, c in [0.3, 0.6]
    st_s = rng.uniform(0.3, 0.9, nG)
    c_s = rng.uniform(0.3, 0.6, nG)
    P0_s = _lower_tri_scatter(rng, nG, st_s, c_s, 0.6, 0.9)
    struct = _mg_dict(st_s, P0_s, (1.0 - c_s) * st_s)
    # moderator: strongly scattering, downscatter-dominant lower-triangular P0
    st_m = np.sort(rng.uniform(0.2, 2.5, nG))
    c_m = rng.uniform(0.97, 0.995, nG)
    P0_m = _lower_tri_scatter(rng, nG, st_m, c_m, 0.3, 0.7)
    moder = _mg_dict(st_m, P0_m, (1.0 - c_m) * st_m)
    return {"fuel": fuel, "structural": struct, "moderator": moder}


def pin_lattice_7g(seed: int = 2, pop: int = 1000000, inactive: int = 50, active: int = 200,
                   transport: str = "transportOperatorDT", tallies: bool = True):
    nG = 7
    mg = pin_lattice_materials(seed, nG)
    lat_map = []
    for row in UO2_MAP:
        for tok in row.split():
            lat_map.append({1: 1, 2: 2, 6: 2}[int(tok)])
    active_t = {}
    if tallies:
        for g in range(1, nG + 1):
            active_t[f"pins_g{g}"] = {
                "type": "collisionClerk",
                "filter": {"type": "energyFilter", "Gtop": g, "Glow": g},
                "map": {"type": "multiMap", "maps": ["xax", "yax"],
                        "xax": {"type": "spaceMap", "axis": "x", "grid": "lin", "min": -10.71, "max": 10.71, "N": 17},
                        "yax": {"type": "spaceMap", "axis": "y", "grid": "lin", "min": -10.71, "max": 10.71, "N": 17}},
                "response": ["flux", "fiss"], "flux": {"type": "fluxResponse"},
                "fiss": {"type": "macroResponse", "MT": -6}}
    inp = {
        "type": "eigenPhysicsPackage", "pop": pop, "active": active, "inactive": inactive,
        "XSdata": "mg", "dataType": "mg", "seed": seed,
        "collisionOperator": {"neutronMG": {"type": "neutronMGstd"}},
        "transportOperator": {"type": transport},
        "inactiveTally": {}, "activeTally": active_t,
        "geometry": {
            "type": "geometryStd", "boundary": [1, 1, 1, 1, 1, 1], "graph": {"type": "shrunk"},
            "surfaces": {"box": {"id": 1, "type": "box", "origin": [0.0, 0.0, 0.0],
                                 "halfwidth": [10.71, 10.71, 10.71]}},
            "cells": {},
            "universes": {
                "root": {"id": 1000, "type": "rootUniverse", "border": 1, "fill": "u<10>"},
                "pin1": {"id": 1, "type": "pinUniverse", "radii": [0.54, 0.0], "fills": ["fuel", "moderator"]},
                "pin2": {"id": 2, "type": "pinUniverse", "radii": [0.54, 0.0], "fills": ["structural", "moderator"]},
                "lat": {"id": 10, "type": "latUniverse", "origin": [0.0, 0.0, 0.0], "pitch": [1.26, 1.26, 0.0],
                        "shape": [17, 17, 0], "padMat": "moderator", "map": lat_map},
            }},
        "nuclearData": {"handles": {"mg": {"type": "baseMgNeutronDatabase", "PN": "P0"}},
                        "materials": {n: {"temp": 300, "composition": {}, "xsFile": n} for n in mg}},
    }
    return inp, mg


CONFIGS = {"inf_medium_2g": inf_medium_2g, "pin_lattice_7g": pin_lattice_7g}
