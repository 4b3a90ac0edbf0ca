"""Flat tables of the C ABI (include/scone_gpu.h) and the compiler that builds them from SCONE
input dictionaries.

The compiler restates the reference's construction code for the multigroup path:

* geometry: csg surfaces/universes (Geometry/csg_class.f90), universe init routines
  (Geometry/Universes/{root,lat,pin}Universe_class.f90 ``init``) and the geometry graph
  (Geometry/geomGraph_class.f90 buildShrunk/buildExtended/setUniqueIDs);
* multigroup data: baseMgNeutronMaterial_class.f90 ``init``, multiScatterMG_class.f90
  ``buildFromDict``, fissionMG_class.f90 ``buildFromDict`` and the majorant of
  baseMgNeutronDatabase_class.f90 ``initMajorant``;
* tallies: tallyAdmin_class.f90 ``init`` (memory addresses), collisionClerk/maps/filters/responses.

All derived floating-point values are computed with the reference's operation order (sequential
sums from 0.0, products as written), so the tables are bit-identical to what SCONE builds.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import dictio

# ---- constants (include/scone_gpu.h) ------------------------------------------------------
MAX_NEST = 8
OUTSIDE_MAT = 0
VOID_MAT = 2147483647
UNDEF_MAT = 2147483646
INF = 2.0 ** 63
SURF_TOL = 1.0e-12
PI = 4.0 * math.atan(1.0)                  # numPrecision.f90 PI
FP_REL_TOL = 1.0e-7
TRANSPORT_ST, TRANSPORT_DT, TRANSPORT_HT = 1, 2, 3
COLL_STD, COLL_IMP = 0, 1
UNI_ROOT, UNI_LAT, UNI_PIN, UNI_CELL = 1, 2, 3, 4
SURF_APLANE, SURF_PLANE, SURF_SPHERE, SURF_CYLINDER, SURF_QUADRIC, SURF_BOX, SURF_TRUNC_CYLINDER = 1, 2, 3, 4, 5, 6, 7
SURF_TRUNC_CONE, SURF_WEDGE = 8, 9
UCELL = -1
U_UNION, U_INTERSECT, U_OPEN, U_CLOSE, U_COMPLEM = 1 << 30, (1 << 30) + 1, (1 << 30) + 2, (1 << 30) + 3, (1 << 30) + 4
OVERLAP_MAT = 2147483645
CLERK_COLLISION, CLERK_TRACK = 0, 1
MAP_SPACE, MAP_MATERIAL, MAP_ENERGY, MAP_CELL = 1, 2, 3, 4
GRID_LIN, GRID_UNSTRUCT = 1, 3
KAPPA_DEFAULT = float(np.float32(202.27))   # fissionMG_class.f90:64 single-precision literal
CHI_TOL = float(np.float32(0.01)) * FP_REL_TOL  # fissionMG_class.f90:268
MACRO_MT = {"total": -1}


# ---- ctypes mirrors of the ABI structs ----------------------------------------------------
class ParticleState(C.Structure):
    _fields_ = [("wgt", C.c_double), ("r", C.c_double * 3), ("dir", C.c_double * 3), ("E", C.c_double),
                ("G", C.c_int32), ("isMG", C.c_int32), ("type", C.c_int32), ("pad0", C.c_int32),
                ("time", C.c_double), ("lambda_", C.c_double), ("matIdx", C.c_int32),
                ("cellIdx", C.c_int32), ("uniqueID", C.c_int32), ("collisionN", C.c_int32),
                ("broodID", C.c_int32), ("pad1", C.c_int32)]


PARTICLE_DTYPE = np.dtype([("wgt", "f8"), ("r", "f8", 3), ("dir", "f8", 3), ("E", "f8"), ("G", "i4"),
                           ("isMG", "i4"), ("type", "i4"), ("pad0", "i4"), ("time", "f8"),
                           ("lambda", "f8"), ("matIdx", "i4"), ("cellIdx", "i4"), ("uniqueID", "i4"),
                           ("collisionN", "i4"), ("broodID", "i4"), ("pad1", "i4")])
assert PARTICLE_DTYPE.itemsize == C.sizeof(ParticleState)


class Box(C.Structure):
    _fields_ = [("n_axes", C.c_int32), ("axes", C.c_int32 * 3), ("bc", C.c_int32 * 6),
                ("origin", C.c_double * 3), ("halfwidth", C.c_double * 3)]


class Universe(C.Structure):
    _fields_ = [("type", C.c_int32), ("n_local", C.c_int32), ("origin", C.c_double * 3),
                ("border", C.c_int32), ("size", C.c_int32 * 3), ("out_local", C.c_int32),
                ("offset_mode", C.c_int32), ("offset_map", C.c_int32), ("pitch", C.c_double * 3),
                ("corner", C.c_double * 3), ("a_bar", C.c_double * 3), ("outline_hw", C.c_double * 3),
                ("n_annuli", C.c_int32), ("annuli", C.c_int32), ("n_cells", C.c_int32), ("cells", C.c_int32),
                ("border_surf", C.c_int32), ("rotated", C.c_int32), ("global_", C.c_int32), ("check_overlap", C.c_int32),
                ("rot", C.c_double * 9)]


class Surface(C.Structure):
    _fields_ = [("type", C.c_int32), ("axis", C.c_int32), ("tol", C.c_double), ("c", C.c_double * 12)]


class MGData(C.Structure):
    _fields_ = [("n_mat", C.c_int32), ("n_groups", C.c_int32), ("xs", C.POINTER(C.c_double)),
                ("P0", C.POINTER(C.c_double)), ("prod", C.POINTER(C.c_double)),
                ("chi", C.POINTER(C.c_double)), ("fissile", C.POINTER(C.c_int32)),
                ("majorant", C.POINTER(C.c_double)), ("collision_xs", C.c_double),
                ("P1", C.POINTER(C.c_double))]


class Map1D(C.Structure):
    _fields_ = [("type", C.c_int32), ("axis", C.c_int32), ("nbins", C.c_int32), ("grid_type", C.c_int32),
                ("data", C.c_int32)]


class Clerk(C.Structure):
    _fields_ = [("gtop", C.c_int32), ("glow", C.c_int32), ("n_maps", C.c_int32), ("handle_virtual", C.c_int32),
                ("maps", Map1D * 3), ("multi", C.c_int32 * 3), ("width", C.c_int32), ("resp", C.c_int32 * 8),
                ("kind", C.c_int32), ("mem_addr", C.c_int64), ("size", C.c_int64)]


class Problem(C.Structure):
    _fields_ = [("transport", C.c_int32), ("st_cache", C.c_int32), ("n_universes", C.c_int32),
                ("root_uni", C.c_int32), ("universes", C.POINTER(Universe)), ("n_boxes", C.c_int32),
                ("border_box", C.c_int32), ("boxes", C.POINTER(Box)), ("graph_size", C.c_int32),
                ("graph_fill", C.POINTER(C.c_int32)), ("graph_id", C.POINTER(C.c_int32)),
                ("n_real", C.c_int32), ("real_data", C.POINTER(C.c_double)), ("n_int", C.c_int32),
                ("int_data", C.POINTER(C.c_int32)), ("mg", MGData), ("n_clerks", C.c_int32),
                ("clerks", C.POINTER(Clerk)), ("n_bins", C.c_int64), ("norm_bin", C.c_int64),
                ("norm_value", C.c_double), ("ht_cutoff", C.c_double), ("n_surfaces", C.c_int32),
                ("surfaces", C.POINTER(Surface)), ("collision", C.c_int32), ("max_split", C.c_int32),
                ("ww_n", C.c_int32), ("fixed_point_tally", C.c_int32), ("ww_map", C.POINTER(Clerk)),
                ("ww_lower", C.POINTER(C.c_double)), ("ww_upper", C.POINTER(C.c_double)),
                ("ww_survival", C.c_double)]


class EigenSettings(C.Structure):
    _fields_ = [("pop", C.c_int32), ("n_inactive", C.c_int32), ("n_active", C.c_int32),
                ("inactive_tally", C.c_int32), ("seed", C.c_int64), ("keff0", C.c_double),
                ("events_per_launch", C.c_int32), ("block_size", C.c_int32), ("cycle_graph", C.c_int32),
                ("refill", C.c_int32), ("select_window", C.c_double)]


class EigenResult(C.Structure):
    _fields_ = [("k_analog", C.c_double * 2), ("k_implicit", C.c_double * 2), ("imp", (C.c_double * 2) * 4),
                ("k_history", C.POINTER(C.c_double)), ("cycle_time", C.POINTER(C.c_double)),
                ("bins_active", C.POINTER(C.c_double)),
                ("bins_inactive", C.POINTER(C.c_double)), ("transport_time", C.c_double),
                ("n_collisions", C.c_int64), ("n_events", C.c_int64), ("n_histories", C.c_int64)]


SRC_POINT = 1
SRC_MATERIAL = 2
SLOT_FREE, SLOT_RUNNING, SLOT_DONE = 0, 1, 2


class Source(C.Structure):
    _fields_ = [("type", C.c_int32), ("isotropic", C.c_int32), ("G", C.c_int32), ("mat", C.c_int32),
                ("n_probG", C.c_int32), ("pad", C.c_int32), ("r", C.c_double * 3), ("dir", C.c_double * 3),
                ("bottom", C.c_double * 3), ("top", C.c_double * 3), ("t_low", C.c_double),
                ("t_high", C.c_double), ("probG", C.POINTER(C.c_double))]


class FixedRequest(C.Structure):
    _fields_ = [("problem", C.POINTER(Problem)), ("source", C.POINTER(Source)), ("pop", C.c_int32),
                ("n_batches", C.c_int32), ("seed", C.c_int64), ("stop_bin", C.c_int64),
                ("stop_rel", C.c_double), ("record", C.c_int32), ("buffer", C.c_int32), ("window", C.c_int32),
                ("adapt", C.c_int32)]


class FixedResult(C.Structure):
    _fields_ = [("state", C.c_int32), ("n_batches", C.c_int32), ("n_bins", C.c_int64),
                ("n_histories", C.c_int64), ("n_collisions", C.c_int64), ("n_segments", C.c_int64)]


class BatchSettings(C.Structure):
    _fields_ = [("transport", C.c_int32), ("max_slots", C.c_int32), ("hbm_budget", C.c_int64),
                ("tile", C.c_int32), ("k0", C.c_int32), ("n_launch", C.c_int32), ("use_graph", C.c_int32),
                ("refill", C.c_int32), ("pad0", C.c_int32)]


def _ptr(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


# ---- MG material (baseMgNeutronMaterial_class.f90:186-291) --------------------------------
class MGMaterial:
    """One material's multigroup data exactly as baseMgNeutronMaterial%init builds it."""

    def __init__(self, d: dict, name: str = "", pn: str = "P0"):
        self.name = name
        nG = int(d["numberOfGroups"])
        if nG < 1:
            raise ValueError("Number of groups is invalid")
        self.nG = nG
        self.fissile = "fission" in d
        P0 = [float(x) for x in d["P0"]]
        prod = [float(x) for x in d["scatteringMultiplicity"]]
        if len(P0) != nG * nG or len(prod) != nG * nG:
            raise ValueError("Invalid size of P0 / scatteringMultiplicity")
        # reshape(temp, [nG, nG]): flat index (gin-1)*nG + (gout-1) is P0(gout, gin)
        self.P0 = P0
        self.prod = prod
        self.P1 = None
        if pn == "P1":
            # multiScatterP1MG_class.f90 buildFromDict: P1 = P1 / P0 * 3 where P0 /= 0, else 0
            P1 = [float(x) for x in d["P1"]]
            if len(P1) != nG * nG:
                raise ValueError(f"Invalid size of P1. Expected: {nG * nG} got: {len(P1)}")
            self.P1 = [(P1[i] / P0[i] * 3.0) if P0[i] != 0.0 else 0.0 for i in range(nG * nG)]
        elif pn != "P0":
            raise ValueError(f"scatterKey: {pn} is wrong. Must be P0 or P1")
        scatter = []
        for gin in range(nG):  # sum(P0, 1): sequential over G_out from 0.0
            s = 0.0
            for gout in range(nG):
                s = s + P0[gin * nG + gout]
            scatter.append(s)
        capture = [float(x) for x in d["capture"]]
        if len(capture) != nG:
            raise ValueError("Capture XSs have wrong size")
        data = np.zeros((nG, 6))
        for g in range(nG):
            data[g, 1] = scatter[g]
            data[g, 2] = capture[g]
        self.chi = [0.0] * nG
        if self.fissile:
            fission = [float(x) for x in d["fission"]]
            nu = [float(x) for x in d["nu"]]
            chi = [float(x) for x in d["chi"]]
            if len(fission) != nG or len(nu) != nG or len(chi) != nG:
                raise ValueError("fission data of wrong size")
            if "kappa" in d:
                kappa = [float(x) for x in d["kappa"]]
            else:
                kappa = [KAPPA_DEFAULT] * nG
            s = 0.0
            for c in chi:
                s = s + c
            if abs(s - 1.0) > CHI_TOL:
                chi = [c / s for c in chi]
            self.chi = chi
            for g in range(nG):
                data[g, 3] = fission[g]
                data[g, 4] = nu[g] * fission[g]
                data[g, 5] = kappa[g] * fission[g]
        for g in range(nG):
            t = data[g, 1] + data[g, 2]
            if self.fissile:
                t = t + data[g, 3]
            data[g, 0] = t
        self.data = data


# ---- geometry compiler -------------------------------------------------------------------
_BC_NAMES = {}


def _is_universe_fill(name):
    return isinstance(name, str) and name.startswith("u<")


class Compiled:
    """Holds numpy arrays behind a ctypes Problem (keeps them alive)."""

    def __init__(self):
        self.keep = []
        self.problem = Problem()
        self.mat_names = []
        self.clerk_names = []
        self.clerk_shapes = []
        self.clerk_info = []
        self.nG = 0
        self.bounds = None

    def hold(self, arr):
        self.keep.append(arr)
        return arr


def _surface_box(sd: dict, bc, raw_bounds=False):
    t = sd["type"]
    b = Box()
    b.bc[:] = [int(x) for x in bc] if bc is not None else [0] * 6
    origin = [float(x) for x in sd["origin"]]
    hw = [float(x) for x in sd["halfwidth"]]
    if t == "box":
        b.n_axes = 3
        b.axes[:] = [0, 1, 2]
        b.origin[:] = origin
        b.halfwidth[:] = hw
        bounds = [origin[i] - hw[i] for i in range(3)] + [origin[i] + hw[i] for i in range(3)]
    elif t in ("xSquareCylinder", "ySquareCylinder", "zSquareCylinder"):
        plane = {"x": [1, 2], "y": [0, 2], "z": [0, 1]}[t[0]]
        axis = {"x": 0, "y": 1, "z": 2}[t[0]]
        b.n_axes = 2
        b.axes[:] = plane + [0]
        # squareCylinder keeps origin/halfwidth of its plane (squareCylinder_class.f90 init)
        b.origin[:] = [origin[plane[0]], origin[plane[1]], 0.0]
        b.halfwidth[:] = [hw[plane[0]], hw[plane[1]], 0.0]
        bounds = [0.0] * 6
        for i, ax in enumerate(plane):
            bounds[ax] = b.origin[i] - b.halfwidth[i]
            bounds[3 + ax] = b.origin[i] + b.halfwidth[i]
        bounds[axis] = -INF
        bounds[3 + axis] = INF
    else:
        raise NotImplementedError(f"border surface type {t} is outside the engine's scope")
    if raw_bounds:
        return b, bounds
    return b, _geometry_bounds(bounds)


def _bool(v, key):
    """dictionary%get of a logical: the integer 1 or 0 (dictionary_class.f90 getBool)."""
    if isinstance(v, bool):
        return v
    if isinstance(v, (int, np.integer)) and int(v) in (0, 1):
        return bool(v)
    raise ValueError(f"Keyword {key} must be 1 or 0")


def fortran_norm2(v):
    """Fortran NORM2 of a short vector, evaluated by the scaled sum of squares gfortran uses."""
    result, scale = 0.0, 1.0
    for x in v:
        if x != 0.0:
            ax = abs(x)
            if scale < ax:
                val = scale / ax
                result = 1.0 + result * val * val
                scale = ax
            else:
                val = ax / scale
                result = result + val * val
    return scale * math.sqrt(result)


def rotation_matrix(phi, theta, psi):
    """genericProcedures.f90:1098-1140 rotationMatrix (Euler angles in degrees), row-major."""
    if phi < 0.0 or phi >= 360.0:
        raise ValueError(f"Angle phi must be in <0;360). Is: {phi}")
    if theta < 0.0 or theta > 180.0:
        raise ValueError(f"Angle theta must be in <0;180>. Is: {theta}")
    if psi < 0.0 or psi >= 360.0:
        raise ValueError(f"Angle psi must be in <0;360). Is: {psi}")
    conv = (2.0 * (4.0 * math.atan(1.0))) / 360.0
    sp, cp = math.sin(phi * conv), math.cos(phi * conv)
    st, ct = math.sin(theta * conv), math.cos(theta * conv)
    ss, cs = math.sin(psi * conv), math.cos(psi * conv)
    return [cs * cp - ct * sp * ss, cs * sp + ct * cp * ss, ss * sp,
            -ss * cp - ct * sp * cs, -ss * sp + ct * cp * cs, cs * st,
            st * sp, -st * cp, ct]


_AXIS = {"x": 0, "y": 1, "z": 2}


def _char_to_int(x):
    """genericProcedures charToInt: the integer an entry spells, or None."""
    if isinstance(x, (int, np.integer)) and not isinstance(x, bool):
        return int(x)
    try:
        return int(str(x).strip())
    except ValueError:
        return None


def union_cell_tokens(definition, cid: int, surf_index):
    """unionCell_class.f90 init, deMorgan and checkPrecedence: the definition ('<', '>', ':', '#'
    and signed surface IDs, adjacent items intersected) -> (simple, tokens), surface tokens being
    +-(surf_index(id) + 1) and operators U_*."""
    ev, senses = [], []          # ev: 1-based occurrence numbers of surfaces, or operator codes
    surface_prev = close_prev = False
    depth = 0
    for item in definition:
        num = _char_to_int(item)
        if num is None:
            e = str(item).strip()
            if e == "<":
                depth += 1
                if surface_prev or close_prev:
                    ev.append(U_INTERSECT)
                ev.append(U_OPEN)
                close_prev = False
            elif e == ">":
                depth -= 1
                ev.append(U_CLOSE)
                close_prev = True
            elif e == ":":
                ev.append(U_UNION)
                close_prev = False
            elif e == "#":
                if surface_prev or close_prev:
                    ev.append(U_INTERSECT)
                ev.append(U_COMPLEM)
                close_prev = False
            else:
                raise ValueError(f"Unrecognised entry in surfaces: {e}")
            surface_prev = False
        else:
            if surface_prev or close_prev:
                ev.append(U_INTERSECT)
            if num == 0:
                raise ValueError(f"Surface ID 0 in cell {cid}")
            k = surf_index(abs(num)) + 1
            senses.append(k if num > 0 else -k)
            ev.append(len(senses))
            surface_prev = True
            close_prev = False
    if not ev:
        raise ValueError(f"Cell {cid} has no surfaces")
    if depth != 0:
        raise ValueError("Inconsistent use of parentheses in cell definition")
    if ev[-1] == U_UNION:
        raise ValueError("Cell definition ends with a union")
    if ev[-1] == U_COMPLEM:
        raise ValueError("Cell definition ends with a complement")
    # deMorgan (unionCell_class.f90): drop each complement and flip what it applies to -- the next
    # surface, or the bracketed group it opens -- in place; the scan then moves on by one
    i = 0
    while i < len(ev) - 1:
        if ev[i] == U_COMPLEM:
            del ev[i]
            end = i
            if ev[i] == U_OPEN:
                d, j = 1, i + 1
                while j < len(ev):
                    if ev[j] == U_OPEN:
                        d += 1
                    elif ev[j] == U_CLOSE:
                        d -= 1
                    if d == 0:
                        break
                    j += 1
                end = min(j, len(ev) - 1)
            for j in range(i, end + 1):
                if ev[j] < U_UNION:
                    senses[ev[j] - 1] = -senses[ev[j] - 1]
                elif ev[j] == U_UNION:
                    ev[j] = U_INTERSECT
                elif ev[j] == U_INTERSECT:
                    ev[j] = U_UNION
        i += 1
    # checkPrecedence: unions and intersections may not mix at one bracket level
    i = 0
    n = len(ev)
    while i < n - 1:
        el1 = ev[i]
        if el1 in (U_UNION, U_INTERSECT):
            j = i + 1
            jumped = False
            while j < n:
                el2 = ev[j]
                if el2 in (U_UNION, U_INTERSECT):
                    if el2 != el1:
                        raise ValueError("The order of operations in a cell is ambiguous")
                    i = j
                    jumped = True
                    break
                if el2 in (U_OPEN, U_CLOSE):
                    i = j + 1
                    jumped = True
                    break
                j += 1
            if jumped:
                continue
        i += 1
    simple = U_UNION not in ev
    if simple and len({abs(x) for x in senses}) != len(senses):
        raise ValueError(f"There are repeated surfaces in definition of cell: {cid}")
    tokens = [senses[e - 1] if e < U_UNION else e for e in ev]
    return int(simple), tokens


def compile_surface(sd: dict, boxes: list | None = None):
    """One surface dictionary (Geometry/Surfaces: the type's init) -> (Surface, bounding box).
    Box and squareCylinder surfaces append their Box to `boxes` and refer to it by index."""
    t = sd["type"]
    S = Surface()
    S.tol = SURF_TOL
    if int(sd["id"]) < 1:
        raise ValueError("Invalid surface id provided. ID must be > 1")
    if t in ("xPlane", "yPlane", "zPlane"):                       # aPlane_class.f90 init
        S.type = SURF_APLANE
        S.axis = _AXIS[t[0]]
        S.c[0] = float(sd[t[0] + "0"])
        bb = [-INF] * 3 + [INF] * 3
        bb[S.axis] = bb[3 + S.axis] = S.c[0]
    elif t == "plane":                                            # plane_class.f90 init
        cf = [float(x) for x in sd["coeffs"]]
        if len(cf) != 4:
            raise ValueError(f"4 plane coefficients must be given. There are: {len(cf)}")
        if all(x == 0.0 for x in cf[:3]):
            raise ValueError("Invalid plane normal. coefficients 1-3 are all 0.0.")
        n = fortran_norm2(cf[:3])
        S.type = SURF_PLANE
        S.c[:4] = [x / n for x in cf]
        bb = [-INF] * 3 + [INF] * 3
    elif t == "sphere":                                           # sphere_class.f90 init
        o = [float(x) for x in sd["origin"]]
        r = float(sd["radius"])
        if len(o) != 3:
            raise ValueError(f"Origin needs to have size 3. Has: {len(o)}")
        if r <= 0.0:
            raise ValueError(f"Radius of sphere must be +ve. Is: {r}")
        S.type = SURF_SPHERE
        S.c[:5] = o + [r, r * r]
        S.tol = 2.0 * r * SURF_TOL
        bb = [o[i] - r for i in range(3)] + [o[i] + r for i in range(3)]
    elif t in ("xCylinder", "yCylinder", "zCylinder"):             # cylinder_class.f90 init/build
        o = [float(x) for x in sd["origin"]]
        r = float(sd["radius"])
        if len(o) != 3:
            raise ValueError(f"Origin needs to have size 3. Has: {len(o)}")
        if r <= 0.0:
            raise ValueError(f"Radius of cylinder must be +ve. Is: {r}")
        S.type = SURF_CYLINDER
        S.axis = _AXIS[t[0]]
        S.c[:5] = o + [r, r * r]
        S.tol = 2.0 * r * SURF_TOL
        bb = [o[i] - r for i in range(3)] + [o[i] + r for i in range(3)]
        bb[S.axis], bb[3 + S.axis] = -INF, INF
    elif t == "quadric":                                          # quadric_class.f90 init
        cf = [float(x) for x in sd["coeffs"]]
        if len(cf) != 10:
            raise ValueError(f"coeffs needs to have size 10. Has: {len(cf)}")
        S.type = SURF_QUADRIC
        S.c[:10] = cf
        S.tol = 2.0 * max(abs(x) for x in cf) * SURF_TOL
        bb = [-INF] * 3 + [INF] * 3
    elif t in ("xTruncCylinder", "yTruncCylinder", "zTruncCylinder"):   # truncCylinder_class.f90 init
        o = [float(x) for x in sd["origin"]]
        r = float(sd["radius"])
        a = float(sd["halfwidth"])
        if len(o) != 3:
            raise ValueError(f"origin must have size 3. Has: {len(o)}")
        if r <= 0.0:
            raise ValueError(f"Radius must be +ve. Is: {r}")
        if a <= 0.0:
            raise ValueError(f"Halfwidth must be +ve. Is: {a}")
        S.type = SURF_TRUNC_CYLINDER
        S.axis = _AXIS[t[0]]
        S.c[:5] = o + [r, a]
        bb = [o[i] - r for i in range(3)] + [o[i] + r for i in range(3)]
        bb[S.axis], bb[3 + S.axis] = o[S.axis] - a, o[S.axis] + a
    elif t in ("xTruncCone", "yTruncCone", "zTruncCone"):         # truncCone_class.f90 init/build
        v = [float(x) for x in sd["vertex"]]
        angle = float(sd["angle"])
        h_min, h_max = float(sd["hMin"]), float(sd["hMax"])
        if len(v) != 3:
            raise ValueError(f"Vertex needs to have size 3. It has size: {len(v)}")
        if angle <= 0.0 or angle >= 90.0:
            raise ValueError("Opening angle of cone must be in the range 0-90 degrees "
                             f"(extremes excluded). It is: {angle}")
        if h_min >= h_max:
            raise ValueError("hMin is greater than or equal to hMax.")
        if math.copysign(abs(h_min), h_max) != h_min and h_min != 0.0 and h_max != 0.0:
            raise ValueError("hMin and hMax have different signs.")
        tangent = math.tan(angle * PI / 180.0)
        S.type = SURF_TRUNC_CONE
        S.axis = _AXIS[t[0]]
        S.c[:6] = v + [tangent * tangent, h_min, h_max]
        S.tol = SURF_TOL * ((abs(h_min) + abs(h_max)) * 0.5 * tangent)
        ax = S.axis
        r_max = max(abs(h_min), abs(h_max)) * math.sqrt(tangent * tangent)
        bb = [v[i] - r_max for i in range(3)] + [v[i] + r_max for i in range(3)]
        bb[ax], bb[3 + ax] = v[ax] + h_min, v[ax] + h_max
    elif t in ("xWedge", "yWedge", "zWedge"):                     # wedge_class.f90 init/build
        o = [float(x) for x in sd["origin"]]
        hw = float(sd["halfwidth"])
        height = float(sd["altitude"])
        if int(sd["id"]) <= 0:
            raise ValueError(f"ID must be <=0. Is: {sd['id']}")
        if len(o) != 3:
            raise ValueError(f"origin must have size 3. Has: {len(o)}")
        if hw <= 0.0:
            raise ValueError("halfwidth cannot be -ve.")
        if height <= 0.0:
            raise ValueError("altitude cannot be -ve.")
        theta = float(sd["opening"])
        phi = float(sd.get("rotation", 0.0))
        if theta <= 0.0 or theta >= 90.0:
            raise ValueError("Opening angle of wedge must be in the range 0-90 degrees "
                             f"(extremes excluded). It is: {theta}")
        theta = theta * PI / 180.0
        if phi < 0.0 or phi > 360.0:
            raise ValueError(f"Rotation angle of wedge must be in the range 0-360 degrees. It is: {phi}")
        phi = phi * PI / 180.0
        n1 = [math.sin(phi - theta), -math.cos(phi - theta)]
        n2 = [-math.sin(phi + theta), math.cos(phi + theta)]
        n3 = [math.cos(phi), math.sin(phi)]
        S.type = SURF_WEDGE
        S.axis = _AXIS[t[0]]
        S.c[:11] = o + [hw, height] + n1 + n2 + n3
        S.tol = SURF_TOL * hw
        ax = S.axis
        p1, p2 = [i for i in range(3) if i != ax]
        tg1 = [-n1[1], n1[0]]
        tg2 = [n2[1], -n2[0]]
        c_t = math.cos(theta)
        q0 = [o[p1], o[p2]]
        q1 = [q0[i] + height * tg1[i] / c_t for i in range(2)]
        q2 = [q0[i] + height * tg2[i] / c_t for i in range(2)]
        bb = [0.0] * 6
        bb[ax], bb[3 + ax] = o[ax] - hw, o[ax] + hw
        bb[p1], bb[p2] = min(q0[0], q1[0], q2[0]), min(q0[1], q1[1], q2[1])
        bb[3 + p1], bb[3 + p2] = max(q0[0], q1[0], q2[0]), max(q0[1], q1[1], q2[1])
    elif t in ("box", "xSquareCylinder", "ySquareCylinder", "zSquareCylinder"):
        if boxes is None:
            raise ValueError("a box surface needs a box table")
        b, bb = _surface_box(sd, None, raw_bounds=True)
        S.type = SURF_BOX
        S.axis = len(boxes)
        boxes.append(b)
    else:
        raise NotImplementedError(f"surface type {t} is outside the engine's scope")
    return S, bb


def _geometry_bounds(bb):
    """geometryStd%bounds: axes that are infinite on both sides become (0, 0)."""
    bb = list(bb)
    for i in range(3):
        if bb[i] <= -INF and bb[i + 3] >= INF:
            bb[i] = 0.0
            bb[i + 3] = 0.0
    return bb


def compile_problem(inp: dict, mg_dicts: dict | None = None, base_dir: str = ".",
                    tally_key: str = "activeTally") -> Compiled:
    """Compile a SCONE input dictionary (eigenPhysicsPackage / fixedSourcePhysicsPackage keys)
    into the flat tables of include/scone_gpu.h.

    mg_dicts: optional {material name: parsed xsFile dictionary} used instead of reading files.
    """
    out = Compiled()
    P = out.problem

    # --- transport operator (transportOperatorFactory_func.f90:36-67)
    top = inp.get("transportOperator", {"type": "transportOperatorST"})
    ttype = top["type"]
    if ttype == "transportOperatorST":
        P.transport = TRANSPORT_ST
    elif ttype == "transportOperatorDT":
        P.transport = TRANSPORT_DT
    elif ttype == "transportOperatorHT":
        # transportOperatorHT_class.f90 init: cutoff default 0.9
        P.transport = TRANSPORT_HT
        P.ht_cutoff = float(top.get("cutoff", 0.9))
    else:
        raise NotImplementedError(f"{ttype} is outside the engine's scope")
    P.st_cache = int(top.get("cache", 1))

    # --- collision processor (collisionProcessorFactory_func.f90:36-72, MG neutrons)
    cop = (inp.get("collisionOperator", {}) or {}).get("neutronMG", {"type": "neutronMGstd"})
    ctype = cop.get("type", "neutronMGstd")
    P.max_split = 1000
    if ctype == "neutronMGstd":
        P.collision = COLL_STD
    elif ctype == "neutronMGimp":
        # neutronMGimp_class.f90 init: maxSplit 1000, weightWindows false
        P.collision = COLL_IMP
        P.max_split = int(cop.get("maxSplit", 1000))
    else:
        raise NotImplementedError(f"collision processor {ctype} is outside the engine's scope")

    # --- materials (materialMenu order = order of the materials dictionary)
    nd = inp["nuclearData"]
    mats_d = nd["materials"]
    names = list(mats_d.keys())
    out.mat_names = names
    mat_idx = {n: i + 1 for i, n in enumerate(names)}
    handles = nd.get("handles", {}).get("mg", {})
    pn = handles.get("PN", "P0")
    mats = []
    for n in names:
        md = mats_d[n]
        if mg_dicts is not None and n in mg_dicts:
            dd = mg_dicts[n]
        else:
            path = md["xsFile"]
            if not os.path.isabs(path):
                path = os.path.join(base_dir, path)
            dd = dictio.read(path)
        mats.append(MGMaterial(dd, n, pn))
    nG = mats[0].nG
    for m in mats:
        if m.nG != nG:
            raise ValueError("Inconsistent # of groups in materials")
    out.nG = nG
    collision_xs = 0.0
    if "avgDist" in handles:
        collision_xs = 1.0 / float(handles["avgDist"])

    # --- geometry
    gd = inp["geometry"]
    bc = gd.get("boundary", [0] * 6)
    surfs = gd.get("surfaces", {})
    surf_by_id = {int(sd["id"]): (name, sd) for name, sd in surfs.items()}
    unis = gd["universes"]
    uni_by_id = {}
    order = []
    for name, ud in unis.items():
        uni_by_id[int(ud["id"])] = ud
        order.append(int(ud["id"]))
    uni_idx = {uid: i + 1 for i, uid in enumerate(order)}
    real_data = []
    int_data = []
    universes = []
    boxes = []
    surfaces = []            # Surface entries, in order of first reference
    surf_index = {}          # surface ID -> 0-based index in surfaces
    cells_d = gd.get("cells", {})
    cell_by_id = {}
    for name, cd in cells_d.items():
        cid = int(cd["id"])
        if cid <= 0:
            raise ValueError(f"Cell ID must be >0. Is {cid}")
        if cid in cell_by_id:
            raise ValueError(f"Cells {name} & {cell_by_id[cid][0]} have the same ID: {cid}")
        cell_by_id[cid] = (name, cd)
    cell_shelf = {cid: k + 1 for k, cid in enumerate(cell_by_id)}   # cellShelf index (definition order)
    fills = {}
    root_uid = None

    def surf_ref(sid):
        if sid not in surf_index:
            if sid not in surf_by_id:
                raise ValueError(f"Surface ID {sid} is not defined")
            S, _ = compile_surface(surf_by_id[sid][1], boxes)
            surf_index[sid] = len(surfaces)
            surfaces.append(S)
        return surf_index[sid]

    def cell_fill(cd):
        # cellShelf_class.f90 init: filltype outside / mat / uni
        ft = cd["filltype"]
        if ft == "outside":
            return OUTSIDE_MAT
        if ft == "mat":
            return fill_of(cd["material"])
        if ft == "uni":
            u = int(cd["universe"])
            if u <= 0:
                raise ValueError(f"Universe ID must be +ve. Is :{u}")
            return -u
        raise ValueError(f"Unknown type of cell filling: {ft}")

    def fill_of(name):
        # universe_inter.f90 charToFill: 'u<id>' -> -id, material name -> matIdx
        if _is_universe_fill(name):
            return -int(name[2:name.rindex(">")])
        if name == "void":
            return VOID_MAT
        if name not in mat_idx:
            raise ValueError(f"Unknown material: {name}")
        return mat_idx[name]

    for uid in order:
        ud = uni_by_id[uid]
        U = Universe()
        U.border_surf = -1
        # universe_inter.f90 setupBase: origin, rotation, global
        org = [float(x) for x in ud.get("origin", [0.0, 0.0, 0.0])]
        if len(org) != 3:
            raise ValueError(f"Origin must have size 3. Has: {len(org)}")
        rot = [float(x) for x in ud.get("rotation", [0.0, 0.0, 0.0])]
        if len(rot) != 3:
            raise ValueError(f"3 rotation angles must be given. Has only: {len(rot)}")
        if any(x != 0.0 for x in rot):
            U.rotated = 1
            U.rot[:] = rotation_matrix(*rot)
        U.global_ = int(_bool(ud.get("global", 0), "global"))
        U.origin[:] = org
        t = ud["type"]
        if t == "rootUniverse":
            if U.rotated:
                raise NotImplementedError("a rotated root universe is outside the engine's scope")
            root_uid = uid
            U.type = UNI_ROOT
            sid = int(ud["border"])
            if sid not in surf_by_id:
                raise ValueError(f"Surface ID {sid} is not defined")
            bsd = surf_by_id[sid][1]
            if bsd["type"] in ("box", "xSquareCylinder", "ySquareCylinder", "zSquareCylinder"):
                b, bounds = _surface_box(bsd, bc)
                U.border = len(boxes)
                boxes.append(b)
            else:
                if bsd["type"] in ("xWedge", "yWedge", "zWedge") and any(int(x) != 0 for x in bc):
                    raise NotImplementedError("reflective and periodic wedge boundaries (wedge_class.f90 "
                                              "explicitBC/transformBC) are outside the engine's scope")
                # surface_inter.f90 setBC: other surfaces take vacuum boundaries only
                if int(bc[0]) != 0:
                    raise ValueError(f"{bsd['type']} supports only VACUUM BCs. Was given: {int(bc[0])}")
                U.border = -1
                U.border_surf = surf_ref(sid)
                bounds = _geometry_bounds(compile_surface(bsd, [])[1])
            out.bounds = bounds
            U.n_local = 2
            fills[uid] = [fill_of(ud["fill"]), OUTSIDE_MAT]
        elif t == "cellUniverse":
            # cellUniverse_class.f90 init: local IDs in the order of 'cells', then UNDEF, OVERLAP
            U.type = UNI_CELL
            ids = [int(x) for x in ud["cells"]]
            U.n_cells = len(ids)
            U.n_local = len(ids) + 2
            U.check_overlap = int(_bool(ud.get("checkOverlap", 0), "checkOverlap"))
            U.cells = len(int_data)
            int_data.extend([0] * (2 * len(ids)))     # record offsets, then cellShelf indices
            f = []
            for i, cid in enumerate(ids):
                if cid not in cell_by_id:
                    raise ValueError(f"Cell ID {cid} is not defined")
                cname, cd = cell_by_id[cid]
                int_data[U.cells + len(ids) + i] = cell_shelf[cid]
                if cd["type"] == "unionCell":
                    simple, tokens = union_cell_tokens(cd["surfaces"], cid, surf_ref)
                    int_data[U.cells + i] = len(int_data)
                    int_data.extend([UCELL, simple, len(tokens)] + tokens)
                    f.append(cell_fill(cd))
                    continue
                if cd["type"] != "simpleCell":
                    raise NotImplementedError(f"cell type {cd['type']} is outside the engine's scope")
                sids = [int(x) for x in cd["surfaces"]]
                if 0 in sids:
                    raise ValueError(f"Surface ID 0 in cell {cid}")
                if len({abs(x) for x in sids}) != len(sids):
                    raise ValueError(f"There are repeated surfaces in definition of cell: {cid}")
                int_data[U.cells + i] = len(int_data)
                int_data.append(len(sids))
                for sid in sids:
                    k = surf_ref(abs(sid)) + 1
                    int_data.append(k if sid > 0 else -k)
                f.append(cell_fill(cd))
            fills[uid] = f + [UNDEF_MAT, OVERLAP_MAT]
        elif t == "latUniverse":
            U.type = UNI_LAT
            pitch = [float(x) for x in ud["pitch"]]
            shape = [int(x) for x in ud["shape"]]
            if shape[2] == 0:
                shape[2] = 1
                pitch[2] = 2.0 * INF
            sizeN = shape
            a_bar = [pitch[i] * 0.5 - pitch[i] * SURF_TOL for i in range(3)]
            corner = [-(float(sizeN[i]) * 0.5 * pitch[i]) for i in range(3)]
            U.size[:] = sizeN
            U.pitch[:] = pitch
            U.a_bar[:] = a_bar
            U.corner[:] = corner
            U.outline_hw[:] = [abs(c) for c in corner]
            nloc = sizeN[0] * sizeN[1] * sizeN[2]
            U.out_local = nloc + 1
            U.n_local = nloc + 1
            mp = [int(x) for x in ud["map"]]
            if len(mp) != nloc:
                raise ValueError("Lattice map size not equal to size implied by shape")
            R = sizeN[1] * sizeN[2]
            N1 = sizeN[0]
            f = [0] * (nloc + 1)
            for j in range(R):          # rows flipped: WYSIWYG map (latUniverse_class.f90:200-206)
                src = R - 1 - j
                for i in range(N1):
                    f[i + j * N1] = -mp[src * N1 + i]
            f[nloc] = fill_of(ud["padMat"])
            fills[uid] = f
            offset = ud.get("offset", 1)
            offset = 1 if (offset in (1, "true", "yes", "y", "T")) else 0
            U.offset_mode = offset
            if "offsetMap" in ud:
                om = [int(x) for x in ud["offsetMap"]]
                U.offset_mode = 2
                U.offset_map = len(int_data)
                flat = [0] * (nloc + 1)
                for j in range(R):
                    src = R - 1 - j
                    for i in range(N1):
                        flat[i + j * N1] = om[src * N1 + i]
                int_data.extend(flat)
        elif t == "pinUniverse":
            U.type = UNI_PIN
            radii = [float(x) for x in ud["radii"]]
            fnames = list(ud["fills"])
            if len(radii) != len(fnames):
                raise ValueError("Size of radii and fills does not match")
            N = len(radii)
            # pinUniverse_class.f90 init: zero radius -> last with INF*1.1, selection sort ascending
            idx = min(range(N), key=lambda k: (radii[k], k))
            if radii[idx] != 0.0:
                raise ValueError("Did not find outermost element with radius 0.0")
            radii[idx], radii[N - 1] = radii[N - 1], radii[idx]
            fnames[idx], fnames[N - 1] = fnames[N - 1], fnames[idx]
            radii[N - 1] = INF * 1.1
            for i in range(N - 2, -1, -1):
                sub = radii[: i + 1]
                m = max(range(i + 1), key=lambda k: (sub[k], -k))
                radii[m], radii[i] = radii[i], radii[m]
                fnames[m], fnames[i] = fnames[i], fnames[m]
            for i in range(N - 1):
                if radii[i] == radii[i + 1]:
                    raise ValueError("Duplicate value of radius")
            U.n_annuli = N
            U.annuli = len(real_data)
            real_data.extend(radii)
            real_data.extend([r * r for r in radii])
            U.n_local = N
            fills[uid] = [fill_of(n) for n in fnames]
        else:
            raise NotImplementedError(f"universe type {t} is outside the engine's scope")
        universes.append(U)
    if root_uid is None:
        raise ValueError("no rootUniverse")

    # universe fills: -ID -> -uniIdx
    idx_fills = {}
    for uid, f in fills.items():
        idx_fills[uni_idx[uid]] = [(-uni_idx[-x] if x < 0 else x) for x in f]

    # geomGraph (geomGraph_class.f90)
    graph_type = gd.get("graph", {"type": "shrunk"}).get("type", "shrunk")
    root_idx = uni_idx[root_uid]
    used = set()

    def collect(u):
        if u in used:
            return
        used.add(u)
        for x in idx_fills[u]:
            if x < 0:
                collect(-x)
    collect(root_idx)
    arr_fill, arr_id = [], []
    if graph_type == "shrunk":
        layed = {root_idx: 1}
        arr_fill.extend(idx_fills[root_idx])
        arr_id.extend([0] * len(idx_fills[root_idx]))
        loc = 0
        while loc < len(arr_fill):
            f = arr_fill[loc]
            if f < 0:
                if -f not in layed:
                    layed[-f] = len(arr_fill) + 1
                    arr_fill.extend(idx_fills[-f])
                    arr_id.extend([0] * len(idx_fills[-f]))
                arr_id[loc] = layed[-f]
            loc += 1
    elif graph_type == "extended":
        arr_fill.extend(idx_fills[root_idx])
        arr_id.extend([0] * len(idx_fills[root_idx]))
        loc = 0
        while loc < len(arr_fill):
            f = arr_fill[loc]
            if f < 0:
                arr_id[loc] = len(arr_fill) + 1
                arr_fill.extend(idx_fills[-f])
                arr_id.extend([0] * len(idx_fills[-f]))
            loc += 1
    else:
        raise ValueError(f"Unknown geometry graph type: {graph_type}")
    cells = 0
    used_mats = set()
    for i, f in enumerate(arr_fill):
        if f > 0:
            cells += 1
            arr_id[i] = cells
            used_mats.add(f)
    active = sorted(m for m in used_mats if m not in (VOID_MAT, UNDEF_MAT, VOID_MAT - 2))

    # nesting depth check
    def depth(u):
        return 1 + max([depth(-x) for x in idx_fills[u] if x < 0] or [0])
    if depth(root_idx) > MAX_NEST:
        raise NotImplementedError(f"nesting deeper than {MAX_NEST} levels")

    # --- MG tables
    xs = np.zeros((len(mats), nG, 6))
    P0 = np.zeros((len(mats), nG * nG))
    prod = np.zeros((len(mats), nG * nG))
    chi = np.zeros((len(mats), nG))
    fiss = np.zeros(len(mats), dtype=np.int32)
    P1 = np.zeros((len(mats), nG * nG)) if pn == "P1" else None
    for i, m in enumerate(mats):
        xs[i] = m.data
        P0[i] = m.P0
        if P1 is not None:
            P1[i] = m.P1
        prod[i] = m.prod
        chi[i] = m.chi
        fiss[i] = 1 if m.fissile else 0
    majorant = np.zeros(nG)
    for g in range(nG):   # initMajorant: max over active materials, densityFactor 1
        x = 0.0
        for mi in active:
            v = xs[mi - 1, g, 0]
            x = v if v > x else x
        majorant[g] = x * 1.0
    out.active_mats = active

    # --- tallies
    if inp.get("type") == "fixedSourcePhysicsPackage" and tally_key == "activeTally":
        tally_key = "tally"
    def compile_map(md, K):
        """A tally map (tallyMapFactory_func.f90: spaceMap, materialMap, multiMap of them) into
        K.maps / K.multi / K.n_maps; returns the bin shape and the axes for printing."""
        if md["type"] == "multiMap":
            subs = [md[n] for n in md["maps"]]
        else:
            subs = [md]
        if len(subs) > 3:
            raise NotImplementedError("multiMap with more than 3 maps")
        K.n_maps = len(subs)
        mul = 1
        shape, axes = [], []
        for k, sd in enumerate(subs):
            M = Map1D()
            if sd["type"] == "spaceMap":
                # spaceMap%init (spaceMap_class.f90:91-134): its keyword checks and messages
                if "grid" not in sd:
                    raise ValueError("Keyword 'grid' must be present")
                if "axis" not in sd:
                    raise ValueError("Keyword 'axis' must be present")
                if sd["axis"] not in ("x", "y", "z"):
                    raise ValueError(f"Unrecognised axis: {sd['axis']} must be x, y or z")
                if sd["grid"] not in ("lin", "unstruct"):
                    raise ValueError("'grid' keyword must be: lin or unstruct")
                M.type = MAP_SPACE
                M.axis = {"x": 0, "y": 1, "z": 2}[sd["axis"]]
                if sd["grid"] == "lin":
                    mn, mx, N = float(sd["min"]), float(sd["max"]), int(sd["N"])
                    M.grid_type = GRID_LIN
                    M.nbins = N
                    M.data = len(real_data)
                    real_data.extend([mn, (mx - mn) / N])
                    edges = [mn] + [mn + (i - 1) * ((mx - mn) / N) for i in range(2, N + 2)]
                else:
                    edges = [float(x) for x in sd["bins"]]
                    M.grid_type = GRID_UNSTRUCT
                    M.nbins = len(edges) - 1
                    M.data = len(real_data)
                    real_data.extend(edges)
                axes.append(("space", sd["axis"], edges))
            elif sd["type"] == "materialMap":
                M.type = MAP_MATERIAL
                mlist = list(sd["materials"])
                undef = sd.get("undefBin", "false") in ("yes", "y", "true", "TRUE", "T")
                nb = len(mlist) + (1 if undef else 0)
                default = len(mlist) + 1 if undef else 0
                table = [default] * (len(mats) + 2)
                for bi, mn in enumerate(mlist):
                    table[mat_idx[mn]] = bi + 1
                M.nbins = nb
                M.data = len(int_data)
                int_data.extend(table)
                axes.append(("material", mlist, undef))
            elif sd["type"] == "energyMap":
                # energyMap_class.f90: bins by energy; a multigroup particle maps to bin 0 (no score)
                grid = sd.get("grid")
                if grid in ("lin", "log"):
                    mn, mx, N = float(sd["min"]), float(sd["max"]), int(sd["N"])
                    if N < 1:
                        raise ValueError("Number of bins must be +ve")
                    if abs((mx - mn) / mx) < 1e-7:
                        raise ValueError("Minimum value must be smaller then maximum above realtive FP tolerance")
                    if grid == "lin":
                        step = (mx - mn) / N
                        edges = [mn] + [mn + (i - 1) * step for i in range(2, N + 2)]
                    else:
                        if mn <= 0.0:
                            raise ValueError("For logarithmic grid minimum must be +ve")
                        step = math.log(mx / mn) / N
                        edges = [mn]
                        for _ in range(N):
                            edges.append(edges[-1] * math.exp(step))
                elif grid == "unstruct":
                    edges = [float(x) for x in sd["bins"]]
                    if len(edges) < 2 or any(edges[i + 1] < edges[i] for i in range(len(edges) - 1)):
                        raise ValueError("Provided grid is not sorted")
                elif grid == "predef":
                    raise NotImplementedError("predefined energy grids (preDefEnergyGrids) are outside the "
                                              "engine's scope")
                else:
                    raise ValueError("'grid' keyword must be: lin, log, usntruct or predef")
                M.type = MAP_ENERGY
                M.nbins = len(edges) - 1
                axes.append(("energy", edges))
            elif sd["type"] == "cellMap":
                # cellMap_class.f90: bin of the cellShelf index of the lowest-level cell
                clist = [int(x) for x in sd["cells"]]
                ub = str(sd.get("undefBin", "false"))
                if ub in ("yes", "y", "true", "TRUE", "T"):
                    undef = True
                elif ub in ("no", "n", "false", "FALSE", "F"):
                    undef = False
                else:
                    raise ValueError(f"{ub} is an unrecognised entry!")
                for c_id in clist:
                    if c_id not in cell_shelf:
                        raise ValueError(f"Cell ID {c_id} is not defined")
                default = len(clist) + 1 if undef else 0
                table = [default] * (len(cell_shelf) + 1)
                for bi, c_id in enumerate(clist):
                    table[cell_shelf[c_id]] = bi + 1
                M.type = MAP_CELL
                M.nbins = len(clist) + (1 if undef else 0)
                M.data = len(int_data)
                M.axis = len(cell_shelf)
                int_data.extend(table)
                axes.append(("cell", clist, undef))
            else:
                raise NotImplementedError(f"tally map {sd['type']} is outside the engine's scope")
            K.maps[k] = M
            K.multi[k] = mul
            mul *= M.nbins
            shape.append(M.nbins)
        return shape, axes

    # --- weight windows (neutronMGimp 'weightWindows', weightWindowsField_class.f90 init)
    ww = None
    if P.collision == COLL_IMP and _bool(cop.get("weightWindows", 0), "weightWindows"):
        if inp.get("type") == "eigenPhysicsPackage":
            raise NotImplementedError("weight windows in eigenvalue runs are outside the engine's scope")
        vr = inp.get("varianceReduction")
        if not vr or vr.get("type") != "weightWindowsField":
            raise ValueError("weightWindows needs 'varianceReduction { type weightWindowsField; file ...; }'")
        path = vr["file"]
        if not os.path.isabs(path):
            path = os.path.join(base_dir, path)
        wd = dictio.read(path)
        K = Clerk()
        shape, _ = compile_map(wd["map"], K)
        n = int(np.prod(shape))
        lower = np.ascontiguousarray([float(x) for x in wd["wLower"]], dtype=np.float64)
        upper = np.ascontiguousarray([float(x) for x in wd["wUpper"]], dtype=np.float64)
        surv = float(wd["constSurvival"])
        if len(lower) != n or len(upper) != n:
            raise ValueError("Size of weight window net does not match the map provided")
        if surv <= 1.0:
            raise ValueError("Survival constant must be bigger than one")
        if np.any(lower * surv > upper):
            raise ValueError("Russian Roulette survival weight must be smaller than the weight window upper weight")
        if np.any(lower < 0.0) or np.any(upper < 0.0):
            raise ValueError("Some weights in the weight windows provided are negative")
        ww = (K, n, lower, upper, surv)

    tal = inp.get(tally_key, {}) or {}
    clerks = []
    addr = 0
    norm_name = tal.get("norm")
    norm_bin = -1
    for cname, cd in tal.items():
        if not isinstance(cd, dict):
            continue
        ctype = cd.get("type")
        if ctype not in ("collisionClerk", "trackClerk"):
            raise NotImplementedError(f"tally clerk {ctype} is outside the engine's scope")
        K = Clerk()
        K.handle_virtual = 1
        if ctype == "trackClerk":   # trackClerk_class.f90: path reports only
            K.kind = CLERK_TRACK
        else:
            hv = cd.get("handleVirtual", 1)
            K.handle_virtual = 0 if hv in (0, "false", "no", "n", "F") else 1
        if "filter" in cd:
            fd = cd["filter"]
            if fd.get("type") != "energyFilter" or "Gtop" not in fd:
                raise NotImplementedError("only MG energyFilter filters are supported")
            K.gtop, K.glow = int(fd["Gtop"]), int(fd["Glow"])
            if K.gtop > K.glow:
                raise ValueError("Gtop larger than Glow")
        shape, axes = compile_map(cd["map"], K) if "map" in cd else ([], [])
        resp = list(cd["response"])
        if len(resp) > 8:
            raise NotImplementedError("more than 8 responses")
        K.width = len(resp)
        for i, rn in enumerate(resp):
            rd = cd[rn]
            if rd["type"] == "fluxResponse":
                K.resp[i] = 0
            elif rd["type"] == "macroResponse":
                mt = int(rd["MT"])
                pos = {1: -1, 2: -3, 3: -22, 101: -2, 18: -6, 27: -21}
                if mt > 0:
                    mt = pos.get(mt, None)
                    if mt is None:
                        raise NotImplementedError("microscopic MT responses need CE data")
                K.resp[i] = mt
            else:
                raise NotImplementedError(f"response {rd['type']} is outside the engine's scope")
        size = K.width * int(np.prod(shape)) if shape else K.width
        K.mem_addr = addr
        K.size = size
        if norm_name == cname:
            norm_bin = addr
        addr += size
        clerks.append(K)
        out.clerk_names.append(cname)
        out.clerk_shapes.append([K.width] + shape)
        out.clerk_info.append({"name": cname, "kind": ctype, "addr": int(K.mem_addr), "size": int(size), "axes": axes,
                               "multi": "map" in cd and cd["map"]["type"] == "multiMap",
                               "width": K.width, "shape": [K.width] + shape})

    # --- assemble
    uni_arr = (Universe * len(universes))(*universes)
    box_arr = (Box * len(boxes))(*boxes)
    clerk_arr = (Clerk * max(1, len(clerks)))(*clerks)
    out.hold(uni_arr)
    out.hold(box_arr)
    out.hold(clerk_arr)
    gf = out.hold(np.asarray(arr_fill, dtype=np.int32))
    gi = out.hold(np.asarray(arr_id, dtype=np.int32))
    rd = out.hold(np.asarray(real_data + [0.0], dtype=np.float64))
    idt = out.hold(np.asarray(int_data + [0], dtype=np.int32))
    xs_c = out.hold(np.ascontiguousarray(xs.reshape(-1)))
    P0_c = out.hold(np.ascontiguousarray(P0.reshape(-1)))
    prod_c = out.hold(np.ascontiguousarray(prod.reshape(-1)))
    chi_c = out.hold(np.ascontiguousarray(chi.reshape(-1)))
    fiss_c = out.hold(fiss)
    maj_c = out.hold(majorant)
    P.n_universes = len(universes)
    P.root_uni = root_idx
    P.universes = C.cast(uni_arr, C.POINTER(Universe))
    P.n_boxes = len(boxes)
    P.border_box = universes[root_idx - 1].border
    if ww is not None:
        K, n, lower, upper, surv = ww
        ww_arr = (Clerk * 1)(K)
        out.hold(ww_arr)
        P.ww_n = n
        P.ww_map = C.cast(ww_arr, C.POINTER(Clerk))
        P.ww_lower = _ptr(out.hold(lower), C.c_double)
        P.ww_upper = _ptr(out.hold(upper), C.c_double)
        P.ww_survival = surv
    if surfaces:
        surf_arr = (Surface * len(surfaces))(*surfaces)
        out.hold(surf_arr)
        P.n_surfaces = len(surfaces)
        P.surfaces = C.cast(surf_arr, C.POINTER(Surface))
    P.boxes = C.cast(box_arr, C.POINTER(Box))
    P.graph_size = len(arr_fill)
    P.graph_fill = _ptr(gf, C.c_int32)
    P.graph_id = _ptr(gi, C.c_int32)
    P.n_real = len(real_data)
    P.real_data = _ptr(rd, C.c_double)
    P.n_int = len(int_data)
    P.int_data = _ptr(idt, C.c_int32)
    P.mg.n_mat = len(mats)
    P.mg.n_groups = nG
    P.mg.xs = _ptr(xs_c, C.c_double)
    P.mg.P0 = _ptr(P0_c, C.c_double)
    P.mg.prod = _ptr(prod_c, C.c_double)
    if P1 is not None:
        P.mg.P1 = _ptr(out.hold(np.ascontiguousarray(P1.reshape(-1))), C.c_double)
    P.mg.chi = _ptr(chi_c, C.c_double)
    P.mg.fissile = _ptr(fiss_c, C.c_int32)
    P.mg.majorant = _ptr(maj_c, C.c_double)
    P.mg.collision_xs = collision_xs
    P.n_clerks = len(clerks)
    P.clerks = C.cast(clerk_arr, C.POINTER(Clerk))
    P.n_bins = addr
    P.norm_bin = norm_bin
    P.norm_value = float(tal.get("normVal", 1.0))
    # extension of the engine (no SCONE keyword): `fixedPointTally 1;` in the tally dictionary or
    # at the top of the input makes BIN sums order-independent (scone_gpu.h fixed_point_tally)
    P.fixed_point_tally = int(_bool(tal.get("fixedPointTally", inp.get("fixedPointTally", 0)), "fixedPointTally"))
    out.xs = xs
    out.majorant = majorant
    out.mats = mats
    out.graph = (arr_fill, arr_id)
    out.uni_idx = uni_idx
    return out


class CompiledSource:
    """A ctypes Source and the arrays behind it."""

    def __init__(self):
        self.source = Source()
        self.keep = []


def compile_source(sd: dict, comp: Compiled) -> CompiledSource:
    """Compile a fixed-source ``source`` dictionary (sourceFactory: pointSource_class.f90:73-158,
    materialSource_class.f90:82-148) against the compiled problem ``comp``.  Values stay as
    written; the engine normalises ``dir`` and ``probG`` as pointSource%init does."""
    out = CompiledSource()
    S = out.source
    t = sd["type"]
    if t == "pointSource":
        if sd.get("particle", "neutron") != "neutron":
            raise NotImplementedError("photon sources are outside the engine's scope")
        S.type = SRC_POINT
        r = [float(x) for x in sd["r"]]
        if len(r) != 3:
            raise ValueError("Source position must have three components")
        S.r[:] = r
        S.isotropic = 0 if "dir" in sd else 1
        if "dir" in sd:
            d = [float(x) for x in sd["dir"]]
            if len(d) != 3:
                raise ValueError("Source direction must have three components")
            S.dir[:] = d
        if "E" in sd:
            raise NotImplementedError("continuous-energy sources are outside the engine's scope")
        if "probG" in sd and "G" in sd:
            raise ValueError("Source may be either monoenergetic or a distribution, not both")
        if "probG" in sd:
            pg = np.asarray([float(x) for x in sd["probG"]], dtype=np.float64)
            if pg.size != comp.nG:
                raise ValueError(f"Source energy group distribution must have {comp.nG} groups")
            out.keep.append(pg)
            S.n_probG = pg.size
            S.probG = _ptr(pg, C.c_double)
        elif "G" in sd:
            S.G = int(sd["G"])
        else:
            raise ValueError("Must specify source energy, either Energy(E) or Group distribution (probG)")
    elif t == "materialSource":
        if sd.get("data", "ce") != "mg":
            raise ValueError("materialSource of an MG problem needs 'data mg;'")
        S.type = SRC_MATERIAL
        S.G = int(sd.get("G", 1))
        name = sd["mat"]
        if name not in comp.mat_names:
            raise ValueError(f"Source material {name} was not found in the material definitions")
        S.mat = comp.mat_names.index(name) + 1
        if "boundingBox" in sd:
            bb = [float(x) for x in sd["boundingBox"]]
            if len(bb) != 6:
                raise ValueError("Bounding box must have 6 entries")
        else:
            bb = list(comp.bounds)
        S.bottom[:] = bb[0:3]
        S.top[:] = bb[3:6]
        if "boundingTime" in sd:
            bt = [float(x) for x in sd["boundingTime"]]
            if len(bt) != 2:
                raise ValueError("Bounding time must have 2 entries")
            if bt[1] < bt[0]:
                raise ValueError("tHigh is less than tLow")
            S.t_low, S.t_high = bt
    else:
        raise NotImplementedError(f"source type {t} is outside the engine's scope")
    if S.G < 0 or S.G > comp.nG:
        raise ValueError("source group outside 1..nG")
    return out


def bank_array(n: int) -> np.ndarray:
    return np.zeros(n, dtype=PARTICLE_DTYPE)


def bank_ptr(bank: np.ndarray):
    return bank.ctypes.data_as(C.POINTER(ParticleState))
