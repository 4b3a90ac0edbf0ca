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
* ``core_3d``         -- pin -> 17x17 assembly -> 15x15 core lattice with a 20 cm reflector ring,
                         400 cm tall vacuum box, 70 groups, 200x200x50 mesh x 70 group flux tally
                         (1.4e8 bins).

Fixed-source problems (fixedSourcePhysicsPackage dictionaries, ``source`` and ``tally`` keys):

* ``inf_medium_fixed`` -- non-multiplying infinite medium (reflective cube) with an isotropic
                         point source of group distribution probG; the collision-estimator flux per
                         source neutron is (T - S)^-1 probG.
* ``slab_fixed``       -- 1D slab (vacuum in x, reflective in y and z) of several regions with a
                         material (volume) source and a mesh flux tally.
* ``batch_problem``    -- problem p of the 512-problem batch: 1D-3D boxes, 1-8 regions, 2-70
                         groups, 1e3-1e6 histories in 10 batches, point or volume source, mesh flux
                         tally of 1e3-1e5 bins; seeds from splitmix64(0xC3 ^ p).
* ``lattice_request``  -- a fixed-source solve on the 17x17 pin-lattice problem (fission chains
                         followed as secondaries), histories log-uniform in 1e3-1e6, with a
                         total-flux bin for early stopping; the serving workload.
* ``inf_medium_fixed_fissile`` -- the 2-group infinite medium at k_inf < 1 with a point source;
                         flux per source neutron (T - S - chi nuSf^T)^-1 q.
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


def splitmix64(x: int) -> int:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) of x + golden gamma."""
    m = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def _scatter_matrix(rng, nG, sig_t, c):
    """Full P0[g_out, g_in] with downscatter-dominant shape and a little upscatter, column g_in
    summing to c[g_in] * sig_t[g_in]."""
    P0 = np.zeros((nG, nG))
    for g in range(nG):
        w = np.zeros(nG)
        for go in range(nG):
            d = go - g
            w[go] = rng.uniform(0.2, 1.0) * (np.exp(-0.7 * d) if d >= 0 else 0.05 * np.exp(1.5 * d))
        P0[:, g] = w / w.sum() * c[g] * sig_t[g]
    return P0


def fixed_material(rng, nG, c_lo=0.3, c_hi=0.95, st_lo=0.2, st_hi=2.0):
    sig_t = rng.uniform(st_lo, st_hi, nG)
    c = rng.uniform(c_lo, c_hi, nG)
    P0 = _scatter_matrix(rng, nG, sig_t, c)
    return _mg_dict(sig_t, P0, (1.0 - c) * sig_t)


def _fixed_inp(geometry, materials, source, tally, pop, cycles, seed, transport):
    return {
        "type": "fixedSourcePhysicsPackage", "pop": pop, "cycles": cycles, "XSdata": "mg", "dataType": "mg",
        "seed": seed,
        "collisionOperator": {"neutronMG": {"type": "neutronMGstd"}},
        "transportOperator": {"type": transport},
        "source": source, "tally": tally, "geometry": geometry,
        "nuclearData": {"handles": {"mg": {"type": "baseMgNeutronDatabase", "PN": "P0"}},
                        "materials": {n: {"temp": 300, "composition": {}, "xsFile": n} for n in materials}},
    }


def _flux_clerk(gtop=None):
    d = {"type": "collisionClerk", "response": ["flux"], "flux": {"type": "fluxResponse"}}
    if gtop is not None:
        d["filter"] = {"type": "energyFilter", "Gtop": gtop, "Glow": gtop}
    return d


def inf_medium_fixed(seed: int = 3, nG: int = 3, pop: int = 2000, cycles: int = 20,
                     transport: str = "transportOperatorST"):
    rng = np.random.default_rng(seed)
    mg = fixed_material(rng, nG, 0.5, 0.9, 0.5, 2.0)
    probG = rng.uniform(0.1, 1.0, nG)
    tally = {"total": _flux_clerk()}
    for g in range(1, nG + 1):
        tally[f"flux_g{g}"] = _flux_clerk(g)
    geom = {"type": "geometryStd", "boundary": [1, 1, 1, 1, 1, 1], "graph": {"type": "shrunk"},
            "surfaces": {"cube": {"id": 1, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [5.0, 5.0, 5.0]}},
            "cells": {},
            "universes": {"root": {"id": 1, "type": "rootUniverse", "border": 1, "fill": "medium"}}}
    src = {"type": "pointSource", "r": [0.0, 0.0, 0.0], "particle": "neutron", "probG": [float(x) for x in probG]}
    return _fixed_inp(geom, {"medium": mg}, src, tally, pop, cycles, seed, transport), {"medium": mg}


def inf_medium_fixed_flux(mg: dict, probG):
    """Expected collision-estimator flux per source neutron: (T - S)^-1 q, q = probG / sum."""
    nG = mg["numberOfGroups"]
    P0 = np.array(mg["P0"]).reshape(nG, nG).T
    tot = P0.sum(axis=0) + np.array(mg["capture"])
    q = np.asarray(probG, dtype=float)
    q = q / q.sum()
    return np.linalg.solve(np.diag(tot) - P0, q)


def inf_medium_fixed_fissile(k_inf: float = 0.6, pop: int = 2000, cycles: int = 20,
                             transport: str = "transportOperatorST", seed: int = 1):
    """The 2-group infinite medium with nu scaled to k_inf < 1, driven by a fixed point source
    (probG): secondaries from fission are transported in the source neutron's history."""
    inp, mg = inf_medium_2g(seed=seed, transport=transport)
    m = dict(mg["medium"])
    m["nu"] = [x * k_inf for x in m["nu"]]
    tally = {"total": _flux_clerk()}
    for g in (1, 2):
        tally[f"flux_g{g}"] = _flux_clerk(g)
    src = {"type": "pointSource", "r": [0.0, 0.0, 0.0], "particle": "neutron", "probG": [0.3, 0.7]}
    out = _fixed_inp(inp["geometry"], {"medium": m}, src, tally, pop, cycles, seed, transport)
    return out, {"medium": m}


def fixed_flux_with_fission(mg: dict, probG):
    """Collision-estimator flux per source neutron with fission secondaries (k_eff = 1 in the
    site count): (T - S - chi nuSf^T)^-1 q."""
    nG = mg["numberOfGroups"]
    P0 = np.array(mg["P0"]).reshape(nG, nG).T
    fis = np.array(mg.get("fission", [0.0] * nG))
    nu = np.array(mg.get("nu", [0.0] * nG))
    chi = np.array(mg.get("chi", [0.0] * nG))
    chi = chi / chi.sum() if chi.sum() > 0 else chi
    tot = P0.sum(axis=0) + np.array(mg["capture"]) + fis
    q = np.asarray(probG, dtype=float)
    q = q / q.sum()
    return np.linalg.solve(np.diag(tot) - P0 - np.outer(chi, nu * fis), q)


def _region_geometry(dim, hw, mats, bc):
    """Box root -> lattice (k,1,1 or k,1,0) -> single-fill pin universe per region."""
    k = len(mats)
    pitch_z = 2.0 * hw[2] if dim == 3 else 0.0
    unis = {"root": {"id": 1000, "type": "rootUniverse", "border": 1, "fill": "u<500>"},
            "lat": {"id": 500, "type": "latUniverse", "origin": [0.0, 0.0, 0.0],
                    "pitch": [2.0 * hw[0] / k, 2.0 * hw[1], pitch_z], "shape": [k, 1, 1 if dim == 3 else 0],
                    "padMat": mats[0], "map": list(range(1, k + 1))}}
    for i, m in enumerate(mats):
        unis[f"reg{i + 1}"] = {"id": i + 1, "type": "pinUniverse", "radii": [0.0], "fills": [m]}
    return {"type": "geometryStd", "boundary": bc, "graph": {"type": "shrunk"},
            "surfaces": {"box": {"id": 1, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": list(hw)}},
            "cells": {}, "universes": unis}


def _mesh_clerk(dim, hw, n_bins):
    axes = ["x", "y", "z"][:dim]
    n_ax = max(1, int(round(n_bins ** (1.0 / dim))))
    maps = {"type": "multiMap", "maps": [f"{a}ax" for a in axes]}
    for i, a in enumerate(axes):
        maps[f"{a}ax"] = {"type": "spaceMap", "axis": a, "grid": "lin", "min": -hw[i], "max": hw[i], "N": n_ax}
    return {"type": "collisionClerk", "map": maps, "response": ["flux"], "flux": {"type": "fluxResponse"}}


def slab_fixed(seed: int = 4, nG: int = 4, regions: int = 3, pop: int = 2000, cycles: int = 10,
               transport: str = "transportOperatorST", n_bins: int = 40, track: bool = False):
    """Fixed source in a 1D slab of `regions` materials.  ``track``: the mesh and a per-material
    absorption tally are also scored by trackClerks (path-length estimators)."""
    rng = np.random.default_rng(seed)
    mats = {f"m{i + 1}": fixed_material(rng, nG) for i in range(regions)}
    hw = [float(rng.uniform(5.0, 20.0)), 10.0, 10.0]
    geom = _region_geometry(1, hw, list(mats), [0, 0, 1, 1, 1, 1])
    tally = {"total": _flux_clerk(), "mesh": _mesh_clerk(1, hw, n_bins)}
    if track:
        tally["track_mesh"] = dict(_mesh_clerk(1, hw, n_bins), type="trackClerk")
        tally["track_mats"] = {"type": "trackClerk", "map": {"type": "materialMap", "materials": list(mats)},
                               "filter": {"type": "energyFilter", "Gtop": 1, "Glow": max(1, nG - 1)},
                               "response": ["flux", "abs"], "flux": {"type": "fluxResponse"},
                               "abs": {"type": "macroResponse", "MT": -2}}
    src = {"type": "materialSource", "data": "mg", "mat": "m2" if regions > 1 else "m1", "G": 1}
    return _fixed_inp(geom, mats, src, tally, pop, cycles, seed, transport), mats


def batch_problem(p: int, transport: str = "transportOperatorDT", scale: float = 1.0):
    """Problem p (0..511) of the 512-problem batch; ``scale`` multiplies the history count (for
    small test versions).  Returns (input dict, mg dicts, histories per batch, batches)."""
    seed = splitmix64(0xC3 ^ p)
    rng = np.random.default_rng(seed)
    dim = int(rng.integers(1, 4))
    hw = [float(x) for x in rng.uniform(5.0, 50.0, 3)]
    k = int(rng.integers(1, 9))
    nG = int(round(np.exp(rng.uniform(np.log(2.0), np.log(70.0)))))
    hist = int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e6)))))
    cycles = 10
    pop = max(1, int(round(hist * scale / cycles)))
    mats = {f"m{i + 1}": fixed_material(rng, nG) for i in range(k)}
    n_bins = int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e5)))))
    bc = {1: [0, 0, 1, 1, 1, 1], 2: [0, 0, 0, 0, 1, 1], 3: [0, 0, 0, 0, 0, 0]}[dim]
    geom = _region_geometry(dim, hw, list(mats), bc)
    tally = {"total": _flux_clerk(), "mesh": _mesh_clerk(dim, hw, n_bins)}
    if rng.uniform() < 0.5:
        probG = rng.uniform(0.0, 1.0, nG) * np.exp(-0.3 * np.arange(nG))
        src = {"type": "pointSource", "r": [0.0, 0.0, 0.0], "particle": "neutron",
               "probG": [float(x) for x in probG]}
    else:
        src = {"type": "materialSource", "data": "mg", "mat": f"m{int(rng.integers(1, k + 1))}", "G": 1}
    inp = _fixed_inp(geom, mats, src, tally, pop, cycles, int(seed & 0x7FFFFFFF), transport)
    return inp, mats, pop, cycles


def lattice_request(i: int, transport: str = "transportOperatorDT", scale: float = 1.0, fissile: bool = True):
    """Serving request i: a fixed-source solve on the 17x17 pin-lattice problem of
    ``pin_lattice_7g`` (its fissile fuel, structural and moderator materials; k_inf 0.87, so
    fission chains are finite and are followed as secondaries), a material source in the fuel,
    histories log-uniform in 1e3-1e6 split in 10 batches, a total-flux bin (bin 0) for early
    stopping and a per-pin flux map.  ``fissile=False`` turns fission into capture.
    Returns (input dict, mg dicts, histories per batch, batches)."""
    rng = np.random.default_rng(splitmix64(0xC4 ^ i))
    inp, mg = pin_lattice_7g(pop=1000, transport=transport, tallies=False)
    mg = {n: dict(d) for n, d in mg.items()}
    if not fissile:
        f = mg["fuel"]
        f["capture"] = [c + x for c, x in zip(f["capture"], f["fission"])]
        for key in ("fission", "nu", "chi"):
            f.pop(key)
    hist = int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e6)))))
    cycles = 10
    pop = max(1, int(round(hist * scale / cycles)))
    tally = {"total": _flux_clerk(),
             "pins": {"type": "collisionClerk",
                      "map": {"type": "multiMap", "maps": ["xax", "yax"],
                              "xax": {"type": "spaceMap", "axis": "x", "grid": "lin", "min": -10.71, "max": 10.71,
                                      "N": 17},
                              "yax": {"type": "spaceMap", "axis": "y", "grid": "lin", "min": -10.71, "max": 10.71,
                                      "N": 17}},
                      "response": ["flux"], "flux": {"type": "fluxResponse"}}}
    src = {"type": "materialSource", "data": "mg", "mat": "fuel", "G": 1}
    out = _fixed_inp(inp["geometry"], mg, src, tally, pop, cycles, int(rng.integers(1, 1 << 31)), transport)
    out["buffer"] = 1000
    return out, mg, pop, cycles


def core_3d(seed: int = 5, nG: int = 70, pop: int = 10_000_000, inactive: int = 30, active: int = 100,
            transport: str = "transportOperatorDT", n_assemblies: int = 15, mesh=(200, 200, 50),
            height: float = 400.0, reflector: float = 20.0, tallies: bool = True):
    """3D nested-universe core: pin (r 0.54) -> 17x17 assembly lattice (pitch 1.26, C5G7 UO2 pin
    map) -> n x n core lattice (pitch 21.42) whose padding is a reflector ring `reflector` cm thick,
    in a vacuum box `height` cm tall (2D lattices, finite z).  nG groups of synthetic P0 data
    from `seed`; per-group mesh flux clerks multiMap(x, y, z) over the box."""
    mats = pin_lattice_materials(seed, nG)
    rng = np.random.default_rng(seed + 1000)
    st_r = np.sort(rng.uniform(0.3, 2.0, nG))
    c_r = rng.uniform(0.95, 0.995, nG)
    mats["reflector"] = _mg_dict(st_r, _lower_tri_scatter(rng, nG, st_r, c_r, 0.4, 0.8), (1.0 - c_r) * st_r)
    lat_map = []
    for row in UO2_MAP:
        for tok in row.split():
            lat_map.append({1: 1, 2: 2, 6: 2}[int(tok)])
    half = n_assemblies * 21.42 / 2.0 + reflector
    hz = height / 2.0
    active_t = {}
    if tallies:
        mx, my, mz = mesh
        for g in range(1, nG + 1):
            active_t[f"mesh_g{g}"] = {
                "type": "collisionClerk",
                "filter": {"type": "energyFilter", "Gtop": g, "Glow": g},
                "map": {"type": "multiMap", "maps": ["xax", "yax", "zax"],
                        "xax": {"type": "spaceMap", "axis": "x", "grid": "lin", "min": -half, "max": half, "N": mx},
                        "yax": {"type": "spaceMap", "axis": "y", "grid": "lin", "min": -half, "max": half, "N": my},
                        "zax": {"type": "spaceMap", "axis": "z", "grid": "lin", "min": -hz, "max": hz, "N": mz}},
                "response": ["flux"], "flux": {"type": "fluxResponse"}}
    inp = {
        "type": "eigenPhysicsPackage", "pop": pop, "active": active, "inactive": inactive,
        "XSdata": "mg", "dataType": "mg", "seed": seed,
        "collisionOperator": {"neutronMG": {"type": "neutronMGstd"}},
        "transportOperator": {"type": transport},
        "inactiveTally": {}, "activeTally": active_t,
        "geometry": {
            "type": "geometryStd", "boundary": [0, 0, 0, 0, 0, 0], "graph": {"type": "shrunk"},
            "surfaces": {"box": {"id": 1, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [half, half, hz]}},
            "cells": {},
            "universes": {
                "root": {"id": 1000, "type": "rootUniverse", "border": 1, "fill": "u<100>"},
                "pin1": {"id": 1, "type": "pinUniverse", "radii": [0.54, 0.0], "fills": ["fuel", "moderator"]},
                "pin2": {"id": 2, "type": "pinUniverse", "radii": [0.54, 0.0], "fills": ["structural", "moderator"]},
                "asm": {"id": 10, "type": "latUniverse", "origin": [0.0, 0.0, 0.0], "pitch": [1.26, 1.26, 0.0],
                        "shape": [17, 17, 0], "padMat": "moderator", "map": lat_map},
                "core": {"id": 100, "type": "latUniverse", "origin": [0.0, 0.0, 0.0], "pitch": [21.42, 21.42, 0.0],
                         "shape": [n_assemblies, n_assemblies, 0], "padMat": "reflector",
                         "map": [10] * (n_assemblies * n_assemblies)},
            }},
        "nuclearData": {"handles": {"mg": {"type": "baseMgNeutronDatabase", "PN": "P0"}},
                        "materials": {n: {"temp": 300, "composition": {}, "xsFile": n} for n in mats}},
    }
    return inp, mats


def absorbing_sphere(radius: float = 1.7, sig: float = 0.9, pop: int = 4000, cycles: int = 20,
                     transport: str = "transportOperatorST"):
    """Fixed source in a purely absorbing sphere (the root border, vacuum) with a centred isotropic
    point source; a collision-estimator and a track-length (trackClerk) total-flux bin, both
    (1 - exp(-sig R)) / sig per source neutron."""
    mg = {"m1": {"numberOfGroups": 1, "capture": [sig], "scatteringMultiplicity": [1.0], "P0": [0.0]}}
    geom = {"type": "geometryStd", "boundary": [0, 0, 0, 0, 0, 0], "graph": {"type": "shrunk"},
            "surfaces": {"ball": {"id": 1, "type": "sphere", "origin": [0.0, 0.0, 0.0], "radius": radius}},
            "cells": {},
            "universes": {"root": {"id": 1, "type": "rootUniverse", "border": 1, "fill": "m1"}}}
    src = {"type": "pointSource", "r": [0.0, 0.0, 0.0], "G": 1}
    tally = {"total": _flux_clerk(), "track": dict(_flux_clerk(), type="trackClerk")}
    return _fixed_inp(geom, mg, src, tally, pop, cycles, 3, transport), mg


def cell_lattice(seed: int = 2, pop: int = 20000, inactive: int = 5, active: int = 10,
                 transport: str = "transportOperatorST", n: int = 4, fixed: bool = False, outer: bool = False,
                 track: bool = True, shapes: bool = False):
    """Pin lattice built from CSG cells: an n x n lattice (pitch 1.26, reflective box with finite z)
    of two cellUniverse pins made of every surface type the engine reads -- zCylinder and quadric
    annuli with a zTruncCylinder insert (pin A); a square fuel block of x/yPlanes cut by a
    diagonal plane, with a sphere inclusion, in a universe rotated 30 degrees about z (pin B) --
    the 7-group materials of ``pin_lattice_7g``, and a mesh flux tally.  ``fixed``: a fixed-source
    solve with a material source in the fuel instead of an eigenvalue run.  ``outer``: the lattice
    sits in a 1 x 1 lattice of its own (4 nesting levels) entered through a 'global' universe.
    ``track``: a trackClerk scores the per-material flux and fission rate as well.  ``shapes``: a
    third pin (C) joins the lattice, a fuel zTruncCone beside a structural zWedge in moderator, and
    a fourth (D) of two unionCells: fuel in a cylinder or a sphere, and its complement."""
    mg = pin_lattice_materials(seed, 7)
    surfs = {
        "box": {"id": 1, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [0.63 * n, 0.63 * n, 2.0]},
        "cyl": {"id": 2, "type": "zCylinder", "origin": [0.0, 0.0, 0.0], "radius": 0.41},
        "quad": {"id": 3, "type": "quadric", "coeffs": [1.0, 1.0, 0.0, 0, 0, 0, 0, 0, 0, -0.2304]},   # r 0.48
        "rod": {"id": 4, "type": "zTruncCylinder", "origin": [0.0, 0.0, 0.5], "radius": 0.15, "halfwidth": 1.0},
        "xl": {"id": 5, "type": "xPlane", "x0": -0.35}, "xh": {"id": 6, "type": "xPlane", "x0": 0.35},
        "yl": {"id": 7, "type": "yPlane", "y0": -0.35}, "yh": {"id": 8, "type": "yPlane", "y0": 0.35},
        "diag": {"id": 9, "type": "plane", "coeffs": [1.0, 1.0, 0.0, 0.4]},
        "ball": {"id": 10, "type": "sphere", "origin": [0.1, -0.1, 0.3], "radius": 0.12},
    }
    cells = {
        "a_rod": {"id": 1, "type": "simpleCell", "surfaces": [-4], "filltype": "mat", "material": "structural"},
        "a_fuel": {"id": 2, "type": "simpleCell", "surfaces": [4, -2], "filltype": "mat", "material": "fuel"},
        "a_gap": {"id": 3, "type": "simpleCell", "surfaces": [2, -3], "filltype": "mat", "material": "structural"},
        "a_mod": {"id": 4, "type": "simpleCell", "surfaces": [3], "filltype": "mat", "material": "moderator"},
        "b_ball": {"id": 5, "type": "simpleCell", "surfaces": [-10], "filltype": "mat", "material": "structural"},
        "b_fuel": {"id": 6, "type": "simpleCell", "surfaces": [10, 5, -6, 7, -8, -9], "filltype": "mat",
                   "material": "fuel"},
        "b_cut": {"id": 7, "type": "simpleCell", "surfaces": [10, 5, -6, 7, -8, 9], "filltype": "mat",
                  "material": "moderator"},
        "b_w": {"id": 8, "type": "simpleCell", "surfaces": [10, -5], "filltype": "mat", "material": "moderator"},
        "b_e": {"id": 9, "type": "simpleCell", "surfaces": [10, 6], "filltype": "mat", "material": "moderator"},
        "b_s": {"id": 10, "type": "simpleCell", "surfaces": [10, 5, -6, -7], "filltype": "mat", "material": "moderator"},
        "b_n": {"id": 11, "type": "simpleCell", "surfaces": [10, 5, -6, 8], "filltype": "mat", "material": "moderator"},
    }
    if shapes:
        surfs = dict(surfs)
        surfs["cone"] = {"id": 14, "type": "zTruncCone", "vertex": [0.05, -0.03, -1.5], "angle": 10.0,
                         "hMin": 0.5, "hMax": 3.0}
        surfs["wedge"] = {"id": 15, "type": "zWedge", "origin": [-0.1, -0.2, 0.0], "halfwidth": 1.5,
                          "altitude": 0.4, "opening": 25.0, "rotation": 225.0}
        cells = dict(cells)
        cells["c_cone"] = {"id": 20, "type": "simpleCell", "surfaces": [-14], "filltype": "mat", "material": "fuel"}
        cells["c_wedge"] = {"id": 21, "type": "simpleCell", "surfaces": [14, -15], "filltype": "mat",
                            "material": "structural"}
        cells["c_mod"] = {"id": 22, "type": "simpleCell", "surfaces": [14, 15], "filltype": "mat",
                          "material": "moderator"}
        surfs["pinD_cyl"] = {"id": 16, "type": "zCylinder", "origin": [0.2, 0.0, 0.0], "radius": 0.2}
        surfs["pinD_ball"] = {"id": 17, "type": "sphere", "origin": [-0.2, 0.1, 0.3], "radius": 0.25}
        cells["d_in"] = {"id": 23, "type": "unionCell", "surfaces": ["<", "-16", ":", "-17", ">"],
                         "filltype": "mat", "material": "fuel"}
        cells["d_out"] = {"id": 24, "type": "unionCell", "surfaces": ["#", "<", "-16", ":", "-17", ">"],
                          "filltype": "mat", "material": "moderator"}
    rng = np.random.default_rng(seed + 77)
    lat_map = [int(x) for x in rng.choice([21, 22, 23, 24] if shapes else [21, 22], size=n * n)]
    unis = {"root": {"id": 1000, "type": "rootUniverse", "border": 1, "fill": "u<11>" if outer else "u<10>"}}
    if outer:
        unis["outer"] = {"id": 11, "type": "latUniverse", "pitch": [1.26 * n, 1.26 * n, 0.0], "shape": [1, 1, 0],
                         "padMat": "moderator", "map": [12]}
        unis["wrap"] = {"id": 12, "type": "cellUniverse", "global": 1, "cells": [12]}
        cells = dict(cells)
        cells["wrap_all"] = {"id": 12, "type": "simpleCell", "surfaces": [-13], "filltype": "uni", "universe": 10}
        surfs = dict(surfs)
        surfs["big"] = {"id": 13, "type": "sphere", "origin": [0.0, 0.0, 0.0], "radius": 100.0}
    geom = {"type": "geometryStd", "boundary": [1, 1, 1, 1, 1, 1], "graph": {"type": "shrunk"},
            "surfaces": surfs, "cells": cells,
            "universes": {
                **unis,
                "pinA": {"id": 21, "type": "cellUniverse", "cells": [1, 2, 3, 4]},
                "pinB": {"id": 22, "type": "cellUniverse", "rotation": [30.0, 0.0, 0.0],
                         "cells": [5, 6, 7, 8, 9, 10, 11]},
                **({"pinC": {"id": 23, "type": "cellUniverse", "cells": [20, 21, 22]},
                    "pinD": {"id": 24, "type": "cellUniverse", "cells": [23, 24]}} if shapes else {}),
                "lat": {"id": 10, "type": "latUniverse", "pitch": [1.26, 1.26, 0.0], "shape": [n, n, 0],
                        "padMat": "moderator", "map": lat_map}}}
    h = 0.63 * n
    mesh = {"type": "collisionClerk",
            "map": {"type": "multiMap", "maps": ["xax", "yax"],
                    "xax": {"type": "spaceMap", "axis": "x", "grid": "lin", "min": -h, "max": h, "N": 4 * n},
                    "yax": {"type": "spaceMap", "axis": "y", "grid": "lin", "min": -h, "max": h, "N": 4 * n}},
            "response": ["flux", "fiss"], "flux": {"type": "fluxResponse"},
            "fiss": {"type": "macroResponse", "MT": -6}}
    mats = {"type": "materialMap", "materials": ["fuel", "structural", "moderator"]}
    tally = {"mesh": mesh, "mats": {"type": "collisionClerk", "map": mats, "response": ["flux"],
                                    "flux": {"type": "fluxResponse"}}}
    if track:
        tally["track_mats"] = {"type": "trackClerk", "map": mats, "response": ["flux", "fiss"],
                               "flux": {"type": "fluxResponse"}, "fiss": {"type": "macroResponse", "MT": -6}}
    if shapes:
        # per-cell flux (cellMap over every cell, the rest in 'undefined'), the same on paths, and an
        # energyMap clerk (multigroup particles map to no bin)
        cmap = {"type": "cellMap", "cells": [int(cd["id"]) for cd in cells.values()], "undefBin": "true"}
        tally["cells"] = {"type": "collisionClerk", "map": cmap, "response": ["flux"], "flux": {"type": "fluxResponse"}}
        if track:
            tally["track_cells"] = {"type": "trackClerk", "map": cmap, "response": ["flux"],
                                    "flux": {"type": "fluxResponse"}}
        tally["energy"] = {"type": "collisionClerk", "response": ["flux"], "flux": {"type": "fluxResponse"},
                           "map": {"type": "energyMap", "grid": "log", "min": 1e-11, "max": 20.0, "N": 4}}
    nd = {"handles": {"mg": {"type": "baseMgNeutronDatabase", "PN": "P0"}},
          "materials": {m: {"temp": 300, "composition": {}, "xsFile": m} for m in mg}}
    if fixed:
        src = {"type": "materialSource", "data": "mg", "mat": "fuel", "G": 1}
        out = _fixed_inp(geom, mg, src, tally, pop, active, seed, transport)
        out["buffer"] = 200
        return out, mg
    inp = {"type": "eigenPhysicsPackage", "pop": pop, "active": active, "inactive": inactive,
           "XSdata": "mg", "dataType": "mg", "seed": seed,
           "collisionOperator": {"neutronMG": {"type": "neutronMGstd"}},
           "transportOperator": {"type": transport},
           "inactiveTally": {}, "activeTally": tally, "geometry": geom, "nuclearData": nd}
    return inp, mg


def ww_medium(ww_file: str | None, pop: int = 2000, cycles: int = 10, transport: str = "transportOperatorST",
              max_split: int = 1000, buffer: int = 200):
    """Weight windows under neutronMGimp: a 2-group infinite medium (reflective cube, half-width
    4 cm) whose scattering multiplicities grow weights (1.35 in group 1, 1.2 downscatter) and
    shrink them (0.55 in group 2), an isotropic point source in group 1, total and per-group flux
    bins.  ``ww_file``: the weightWindowsField file (None: neutronMGstd, no windows).  The flux
    per source neutron is (T - M.S)^-1 q with M the multiplicities (``ww_medium_flux``)."""
    mg = {"medium": {"numberOfGroups": 2, "capture": [0.5, 0.9],
                     # one line per incident group: P0(g_out, g_in) for g_in = 1, then g_in = 2
                     "P0": [0.4, 0.3, 0.05, 0.6], "scatteringMultiplicity": [1.35, 1.2, 1.0, 0.55]}}
    tally = {"total": _flux_clerk(), "flux_g1": _flux_clerk(1), "flux_g2": _flux_clerk(2)}
    geom = {"type": "geometryStd", "boundary": [1, 1, 1, 1, 1, 1], "graph": {"type": "shrunk"},
            "surfaces": {"cube": {"id": 1, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [4.0, 4.0, 4.0]}},
            "cells": {},
            "universes": {"root": {"id": 1, "type": "rootUniverse", "border": 1, "fill": "medium"}}}
    src = {"type": "pointSource", "r": [0.0, 0.0, 0.0], "particle": "neutron", "G": 1}
    inp = _fixed_inp(geom, mg, src, tally, pop, cycles, 17, transport)
    inp["buffer"] = buffer
    if ww_file is not None:
        inp["collisionOperator"] = {"neutronMG": {"type": "neutronMGimp", "weightWindows": 1,
                                                  "maxSplit": max_split}}
        inp["varianceReduction"] = {"type": "weightWindowsField", "file": ww_file}
    return inp, mg


def ww_medium_flux():
    """Closed-form group fluxes per source neutron of ``ww_medium``: (T - M.S) phi = e_1."""
    P0 = np.array([[0.4, 0.05], [0.3, 0.6]])          # [g_out, g_in]
    mult = np.array([[1.35, 1.0], [1.2, 0.55]])
    tot = np.array([0.5 + 0.7, 0.9 + 0.65])
    return np.linalg.solve(np.diag(tot) - mult * P0, np.array([1.0, 0.0]))


def ww_file_text(lower, upper, survival=2.0, bins=(-4.0, 0.0, 4.0)):
    """A weightWindowsField file (the reference's format): windows over an x spaceMap."""
    b = " ".join(repr(float(x)) for x in bins)
    return (f"map {{ type spaceMap; axis x; grid unstruct; bins ({b}); }}\n"
            f"constSurvival {float(survival)!r};\n"
            f"wLower ({' '.join(repr(float(x)) for x in lower)});\n"
            f"wUpper ({' '.join(repr(float(x)) for x in upper)});\n")


CONFIGS = {"inf_medium_2g": inf_medium_2g, "pin_lattice_7g": pin_lattice_7g, "core_3d": core_3d,
           "cell_lattice": cell_lattice}
