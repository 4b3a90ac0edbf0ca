"""Surfaces, simpleCell and cellUniverse of the oracle, pinned to the reference's unit tests (CPU).

Known answers transcribed from Geometry/Surfaces/Tests/{aPlane,plane,sphere,cylinder,quadric,
truncCylinder,truncCone,wedge}_test.f90, Geometry/Cells/Tests/{simpleCell,unionCell}_test.f90 and
Geometry/Universes/Tests/cellUniverse_test.f90 (tolerances as in those tests), plus closed forms:
rotationMatrix of genericProcedures.f90 against an independent Euler-angle construction, and the
uncollided escape from a purely absorbing sphere.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import binding as O
from helpers import geometry_problem
from scone_amd import configs, tables as T

INF = 2.0 ** 63
SURF_TOL = 1e-12
SQRT2 = math.sqrt(2.0)


def unit(v):
    v = np.asarray(v, dtype=float)
    return list(v / np.linalg.norm(v))


def surf(**sd):
    sd.setdefault("id", 75)
    S, bb = T.compile_surface(sd, [])
    return S, bb


AXES = [0, 1, 2]
PLANE = {0: (1, 2), 1: (0, 2), 2: (0, 1)}


# ---- aPlane_test.f90 ------------------------------------------------------------------------
@pytest.mark.parametrize("ax", AXES)
def test_aplane(ax):
    name = "xyz"[ax]
    S, bb = surf(type=name + "Plane", **{name + "0": 4.3})
    ref = [-INF] * 3 + [INF] * 3
    ref[ax] = ref[3 + ax] = 4.3
    assert bb == ref
    e = [0.0, 0.0, 0.0]
    e[ax] = 1.0
    r = [0.0, 0.0, 0.0]
    r[ax] = 4.3
    mv = lambda r, eps, u: [r[i] + eps * u[i] for i in range(3)]
    assert O.surf_halfspace(S, r, e)
    assert O.surf_halfspace(S, mv(r, -0.5 * SURF_TOL, e), e)
    assert not O.surf_halfspace(S, mv(r, -1.0001 * SURF_TOL, e), e)
    assert O.surf_halfspace(S, mv(r, 2.0, e), e)
    assert not O.surf_halfspace(S, mv(r, -2.0, e), e)
    u2 = list(np.roll(e, 1))
    assert O.surf_halfspace(S, mv(r, 0.5 * SURF_TOL, e), u2)
    assert not O.surf_halfspace(S, mv(r, -0.5 * SURF_TOL, e), u2)
    # distance
    r = [2.0, 2.0, 2.0]
    r[ax] = 0.0
    u = unit([1, 1, 1])
    assert O.surf_distance(S, r, u) == pytest.approx(4.3 * math.sqrt(3.0), rel=1e-7)
    r[ax] = 4.3
    assert O.surf_distance(S, r, u) == INF
    r[ax] = 4.3 - 0.5 * SURF_TOL
    assert O.surf_distance(S, r, u) == INF
    r[ax] = 5.0
    assert O.surf_distance(S, r, u) == INF
    assert O.surf_distance(S, r, [-x for x in u]) == pytest.approx(0.7 * math.sqrt(3.0), rel=1e-7)
    r = [2.0, 2.0, 2.0]
    r[ax] = 0.0
    assert O.surf_distance(S, r, u2) == INF


# ---- plane_test.f90 -------------------------------------------------------------------------
def test_plane():
    S, bb = surf(id=7, type="plane", coeffs=[1.0, 1.0, 1.0, 3.0])
    assert bb == [-INF] * 3 + [INF] * 3
    s3 = math.sqrt(3.0)
    r = [1.0, 1.0, 1.0]
    u = unit([1, 1, 1])
    mv = lambda eps, d: [r[i] + eps * d[i] for i in range(3)]
    assert O.surf_halfspace(S, r, u)
    assert O.surf_halfspace(S, mv(-0.5 * SURF_TOL * s3, u), u)
    assert not O.surf_halfspace(S, mv(-1.00001 * SURF_TOL * s3, u), u)
    assert not O.surf_halfspace(S, mv(-1.0, u), u)
    assert O.surf_halfspace(S, mv(1.0, u), u)
    u2 = unit([-1, 0, 1])
    assert O.surf_halfspace(S, mv(0.5 * SURF_TOL, u), u2)
    assert not O.surf_halfspace(S, mv(-0.5 * SURF_TOL, u), u2)
    r = [-1.0, 0.0, 0.0]
    u = [1.0, 0.0, 0.0]
    assert O.surf_distance(S, r, u) == pytest.approx(4.0, rel=1e-7)
    assert O.surf_distance(S, [3.0, 0.0, 0.0], u) == INF
    assert O.surf_distance(S, [3.0 - SURF_TOL, 0.0, 0.0], u) == INF
    d_ref = s3 * 1.1 * SURF_TOL
    assert O.surf_distance(S, [3.0 - d_ref, 0.0, 0.0], u) == pytest.approx(d_ref, rel=1e-2)
    assert O.surf_distance(S, [4.0, 0.0, 0.0], u) == INF
    assert O.surf_distance(S, [4.0, 0.0, 0.0], [-1.0, 0.0, 0.0]) == pytest.approx(1.0, rel=1e-7)
    assert O.surf_distance(S, [-1.0, 0.0, 0.0], unit([-1, 0, 1])) == INF


def test_plane_coefficients_normalised_by_norm2():
    S, _ = surf(type="plane", coeffs=[3.0, 4.0, 12.0, 26.0])
    assert list(S.c[:4]) == [3.0 / 13.0, 4.0 / 13.0, 12.0 / 13.0, 2.0]
    with pytest.raises(ValueError, match="all 0.0"):
        surf(type="plane", coeffs=[0.0, 0.0, 0.0, 1.0])


# ---- sphere_test.f90 ------------------------------------------------------------------------
def test_sphere():
    S, bb = surf(id=7, type="sphere", origin=[1.0, 2.0, 1.0], radius=2.0)
    assert bb == [-1.0, 0.0, -1.0, 3.0, 4.0, 3.0]
    assert S.tol == 2.0 * 2.0 * SURF_TOL
    r = [-1.0, 2.0, 1.0]
    u = [1.0, 0.0, 0.0]
    mv = lambda eps: [r[i] + eps * u[i] for i in range(3)]
    assert not O.surf_halfspace(S, r, u)
    assert not O.surf_halfspace(S, mv(-0.5 * SURF_TOL), u)
    assert O.surf_halfspace(S, mv(-1.00001 * SURF_TOL), u)
    assert O.surf_halfspace(S, mv(-1.0), u)
    assert not O.surf_halfspace(S, mv(1.0), u)
    assert O.surf_halfspace(S, r, [0.0, 1.0, 0.0])
    r = [-2.0, 2.0, 1.0]
    assert O.surf_distance(S, r, [1.0, 0.0, 0.0]) == pytest.approx(1.0, rel=1e-7)
    ref = 1.5
    u0 = (ref ** 2 + 3.0 ** 2 - 2.0 ** 2) / (2.0 * ref * 3.0)
    u = unit([u0, math.sqrt(1.0 - u0 ** 2), 0.0])
    assert O.surf_distance(S, r, u) == pytest.approx(ref, rel=1e-7)
    r = [-1.0, 2.0, 1.0]
    assert O.surf_distance(S, r, [0.0, 1.0, 0.0]) == INF
    assert O.surf_distance(S, r, [1.0, 0.0, 0.0]) == 4.0
    r2 = [-1.0 - 0.5 * SURF_TOL, 2.0, 1.0]
    assert O.surf_distance(S, r2, [1.0, 0.0, 0.0]) == 4.0 + 0.5 * SURF_TOL
    r = [-0.5, 2.0, 1.0]
    assert O.surf_distance(S, r, [1.0, 0.0, 0.0]) == 3.5
    assert O.surf_distance(S, r, [-1.0, 0.0, 0.0]) == 0.5


# ---- cylinder_test.f90 ----------------------------------------------------------------------
@pytest.mark.parametrize("a", AXES)
def test_cylinder(a):
    S, bb = surf(type="xyz"[a] + "Cylinder", origin=[2.0, 2.0, 2.0], radius=2.0)
    ref = [0.0, 0.0, 0.0, 4.0, 4.0, 4.0]
    ref[a], ref[3 + a] = -INF, INF
    assert bb == ref
    p1, p2 = PLANE[a]

    def vec(va, v1, v2):
        v = [0.0] * 3
        v[a], v[p1], v[p2] = va, v1, v2
        return v

    r = [2.0] * 3
    r[p1] = 0.0
    u = unit(vec(1.0, 1.0, 0.0))
    mv = lambda eps: [r[i] + eps * u[i] for i in range(3)]
    assert not O.surf_halfspace(S, r, u)
    assert not O.surf_halfspace(S, mv(-SURF_TOL), u)
    assert O.surf_halfspace(S, mv(-SQRT2 * 1.00001 * SURF_TOL), u)
    assert O.surf_halfspace(S, mv(-2.0), u)
    assert not O.surf_halfspace(S, mv(2.0), u)
    assert O.surf_halfspace(S, r, unit(vec(1.0, 0.0, 1.0)))
    # distance
    r = [2.0] * 3
    r[p1] = -1.0
    assert O.surf_distance(S, r, unit(vec(1.0, 1.0, 0.0))) == pytest.approx(SQRT2, rel=1e-7)
    u = unit(vec(1.0, math.sqrt(3.0) / 2.0, 0.5))
    assert O.surf_distance(S, r, u) == pytest.approx(1.275200556 * SQRT2, rel=1e-7)
    assert O.surf_distance(S, r, unit(vec(1.0e20, 1.0, 0.0))) == INF
    r = [2.0] * 3
    r[p1] = 0.0
    u = unit(vec(1.0, 1.0, 0.0))
    assert O.surf_distance(S, r, u) == pytest.approx(4.0 * SQRT2, rel=1e-7)
    assert O.surf_distance(S, r, [-x for x in u]) == INF
    assert O.surf_distance(S, r, [-x for x in unit(vec(1.0, 0.0, 1.0))]) == INF
    r[p1] = 0.0 - 0.5 * SURF_TOL
    assert O.surf_distance(S, r, u) == pytest.approx((4.0 + 0.5 * SURF_TOL) * SQRT2, rel=1e-7)
    r = [2.0] * 3
    r[p1] = 1.0
    assert O.surf_distance(S, r, u) == pytest.approx(3.0 * SQRT2, rel=1e-7)
    assert O.surf_distance(S, r, [-x for x in u]) == pytest.approx(SQRT2, rel=1e-7)


# ---- quadric_test.f90 -----------------------------------------------------------------------
def test_quadric():
    cyl, bb = surf(id=7, type="quadric", coeffs=[1.0, 1.0, 0, 0, 0, 0, 0, 0, 0, -4])
    elp, _ = surf(id=3, type="quadric", coeffs=[0.25, 0.0625, 0.0, 0, 0, 0, -0.5, 0, 0, -0.75])
    assert bb == [-INF] * 3 + [INF] * 3
    u = [-1.0, 0.0, 0.0]
    assert not O.surf_halfspace(cyl, [-0.5, 1.0, 10.0], u)
    r = [2.0, 0.0, 10.0]
    mv = lambda r, eps: [r[i] + eps * u[i] for i in range(3)]
    assert not O.surf_halfspace(cyl, r, u)
    assert not O.surf_halfspace(cyl, mv(r, -0.5 * SURF_TOL), u)
    assert O.surf_halfspace(cyl, mv(r, -8.0001 * SURF_TOL), u)
    assert O.surf_halfspace(cyl, mv(r, -1.0), u)
    assert not O.surf_halfspace(cyl, mv(r, 1.0), u)
    assert O.surf_halfspace(cyl, r, [0.0, 1.0, 0.0])
    assert not O.surf_halfspace(elp, [0.0, 2.0, -8.0], u)
    assert O.surf_halfspace(elp, [-1.0, 2.0, -8.0], u)
    r = [3.0, 0.0, -8.0]
    assert not O.surf_halfspace(elp, r, u)
    assert not O.surf_halfspace(elp, mv(r, -0.5 * SURF_TOL), u)
    assert O.surf_halfspace(elp, mv(r, -1.501 * SURF_TOL), u)
    assert O.surf_halfspace(elp, r, [0.0, 1.0, 0.0])
    TOL = 1e-7
    assert O.surf_distance(cyl, [-4.0, 0.0, 1.0], [1.0, 0.0, 0.0]) == pytest.approx(2.0, rel=TOL)
    u = [0.744208408, 0.248069469, 0.620173673]
    assert O.surf_distance(cyl, [-3.0, 0.5, 30.0], u) == pytest.approx(1.634986685, rel=TOL)
    assert O.surf_distance(cyl, [0.0, 2.0, 1.0], [1.0, 0.0, 0.0]) == INF
    u = [0.707106781187, 0.707106781187, 0.0]
    assert O.surf_distance(cyl, [-0.5, -1.0, 1.0], u) == pytest.approx(3.0291621403, rel=TOL)
    r = [-2.0 - 0.5 * SURF_TOL, 0.0, 1.0]
    assert O.surf_distance(cyl, r, [1.0, 0.0, 0.0]) == pytest.approx(4.0 + 0.5 * SURF_TOL, rel=TOL)
    assert O.surf_distance(elp, [0.0, -6.0, 12.0], [0.0, 1.0, 0.0]) == pytest.approx(2.53589838486, rel=TOL)
    assert O.surf_distance(elp, [-1.0, -4.0, 12.0], u) == pytest.approx(1.13137084990, rel=TOL)
    assert O.surf_distance(elp, [0.0, -2.0, 12.0], u) == pytest.approx(4.16282187604, rel=TOL)


# ---- truncCylinder_test.f90 -----------------------------------------------------------------
@pytest.mark.parametrize("ax", AXES)
def test_trunc_cylinder(ax):
    p1, p2 = PLANE[ax]
    origin = [2.0, 2.0, 2.0]
    origin[p1], origin[p2] = 1.0, 2.0
    S, bb = surf(type="xyz"[ax] + "TruncCylinder", origin=origin, halfwidth=1.5, radius=2.0)
    ref = [0.0] * 6
    ref[ax], ref[p1], ref[p2] = 0.5, -1.0, 0.0
    ref[3 + ax], ref[3 + p1], ref[3 + p2] = 3.5, 3.0, 4.0
    assert bb == pytest.approx(ref)

    def pe(va, v1, v2):
        v = [0.0] * 3
        v[ax], v[p1], v[p2] = va, v1, v2
        return v

    u = pe(0.0, 1.0, 0.0)
    for pt, expect in [((5.0, 3.0, 4.0), True), ((1.3, 1.0, 5.0), True), ((1.3, -2.0, 3.3), True),
                       ((3.3, 1.1, 2.2), False), ((2.1, -0.7, 2.1), False)]:
        assert O.surf_halfspace(S, pe(*pt), u) == expect, pt
    r = pe(3.3, -1.0, 2.0)
    u = unit(pe(1.0, 1.0, 0.0))
    assert not O.surf_halfspace(S, r, u)
    assert O.surf_halfspace(S, r, [-x for x in u])
    eps = 0.5 * SURF_TOL
    assert not O.surf_halfspace(S, pe(3.3, -1.0 - eps, 2.0), u)
    assert O.surf_halfspace(S, pe(3.3, -1.0 - 1.001 * SURF_TOL, 2.0), u)
    assert O.surf_halfspace(S, pe(3.3, -1.0 - eps, 2.0), pe(1.0, 0.0, 0.0))
    assert not O.surf_halfspace(S, pe(3.3, -1.0 + eps, 2.0), pe(1.0, 0.0, 0.0))
    r = pe(3.5, 1.3, 1.8)
    u = unit(pe(-1.0, 0.0, 1.0))
    assert not O.surf_halfspace(S, r, u)
    assert O.surf_halfspace(S, r, [-x for x in u])
    u = unit(pe(-1.0, 1.0, 0.0))
    assert not O.surf_halfspace(S, pe(3.5 + eps, 1.3, 1.8), u)
    assert O.surf_halfspace(S, pe(3.5 + 1.001 * SURF_TOL, 1.3, 1.8), u)
    u = unit(pe(0.0, -1.0, 1.0))
    assert O.surf_halfspace(S, pe(3.5 + eps, 1.3, 1.8), u)
    assert not O.surf_halfspace(S, pe(3.5 - eps, 1.3, 1.8), u)
    # distance
    TOL = 1e-7
    r = pe(5.0, 1.0, 2.0)
    assert O.surf_distance(S, r, pe(-1.0, 0.0, 0.0)) == pytest.approx(1.5, rel=TOL)
    assert O.surf_distance(S, r, unit(pe(-1.0, -1.0, 0.0))) == pytest.approx(1.5 * SQRT2, rel=TOL)
    assert O.surf_distance(S, r, unit(pe(-1.0, -2.0, -2.0))) == INF
    assert O.surf_distance(S, r, unit(pe(0.0, -2.0, -2.0))) == INF
    assert O.surf_distance(S, r, unit(pe(1.0, -1.0, -1.0))) == INF
    r = pe(1.5, -2.0, 2.0)
    assert O.surf_distance(S, r, pe(0.0, 1.0, 0.0)) == pytest.approx(1.0, rel=TOL)
    assert O.surf_distance(S, r, unit(pe(1.0, 1.0, 0.0))) == pytest.approx(SQRT2, rel=TOL)
    v = pe(0.0, 1.5, math.sqrt(7.0) * 0.5)
    assert O.surf_distance(S, r, unit(v)) == pytest.approx(float(np.linalg.norm(v)), rel=TOL)
    assert O.surf_distance(S, r, unit(pe(1.0, 1.0, 2.0))) == INF
    assert O.surf_distance(S, r, pe(1.0, 0.0, 0.0)) == INF
    assert O.surf_distance(S, r, pe(1.0, -1.0, 0.0)) == INF
    assert O.surf_distance(S, pe(3.5 + eps, 1.3, 1.8), pe(-1.0, 0.0, 0.0)) == pytest.approx(3.0, rel=TOL)
    assert O.surf_distance(S, pe(3.5 - eps, 1.3, 1.8), pe(1.0, 0.0, 0.0)) == INF


def _v(ax, va, v1, v2):
    """A vector with component va on the axis and v1, v2 on the plane axes (the tests' a, p1, p2)."""
    p1, p2 = PLANE[ax]
    v = [0.0, 0.0, 0.0]
    v[ax], v[p1], v[p2] = va, v1, v2
    return v


def _mv(r, eps, u):
    return [r[i] + eps * u[i] for i in range(3)]


# ---- truncCone_test.f90 ---------------------------------------------------------------------------
@pytest.mark.parametrize("ax", AXES)
def test_trunc_cone(ax):
    S, bb = surf(id=52, type="xyz"[ax] + "TruncCone", vertex=[1.0, 1.0, 1.0], angle=45.0, hMin=1.0, hMax=10.0)
    ref = _v(ax, 2.0, -9.0, -9.0) + _v(ax, 11.0, 11.0, 11.0)
    assert bb == pytest.approx(ref, abs=1e-6)
    assert S.tol == pytest.approx(SURF_TOL * 5.5, abs=1e-6)
    tol = S.tol
    hs = lambda r, u: O.surf_halfspace(S, r, u)
    # testHalfspace
    r, u = _v(ax, 5.0, 5.0, 1.0), unit(_v(ax, 1.0, -1.0, 0.0))
    assert not hs(r, u)
    assert not hs(_mv(r, -0.01 * tol, u), u)
    assert hs(_mv(r, -2.0 * tol, u), u)
    assert hs(_mv(r, -0.5, u), u)
    assert not hs(_mv(r, 0.5, u), u)
    assert hs(r, _v(ax, 0.0, 0.0, -1.0))
    r, u = _v(ax, 11.0, 4.0, 4.0), unit(_v(ax, 1.0, -1.0, -1.0))
    assert hs(r, u)
    assert hs(_mv(r, -0.5 * tol, u), u)
    assert not hs(_mv(r, -2.0 * tol, u), u)
    assert not hs(_mv(r, -0.5, u), u)
    assert hs(_mv(r, 0.5, u), u)
    assert hs(r, _v(ax, 0.0, 0.0, -1.0))
    # testDistance
    d = lambda r, u: O.surf_distance(S, r, u)
    r = _v(ax, 0.0, 4.0, 1.0)
    assert d(r, _v(ax, 1.0, 0.0, 0.0)) == pytest.approx(4.0, rel=1e-7)
    assert d(r, unit(_v(ax, 2.0, -3.0, 0.0))) == pytest.approx(math.sqrt(13.0), rel=1e-7)
    assert d(r, unit(_v(ax, 1.0, 1.0, 0.0))) == INF
    assert d(r, _v(ax, -1.0, 0.0, 0.0)) == INF
    assert d(r, _v(ax, 0.0, 1.0, 0.0)) == INF
    r = _v(ax, 4.0, 4.0, 1.0)
    assert d(r, _v(ax, 0.0, -1.0, 0.0)) == pytest.approx(6.0, rel=1e-7)
    assert d(r, _v(ax, 0.0, 1.0, 0.0)) == INF
    assert d(r, _v(ax, 0.0, 0.0, 1.0)) == INF
    r = _v(ax, 4.0, 4.0 + 0.1 * tol, 1.0)
    assert d(r, _v(ax, 0.0, -1.0, 0.0)) == pytest.approx(6.0 + 0.5 * tol, rel=1e-7)
    r = _v(ax, 4.0 + 0.1 * tol, 4.0, 1.0)
    assert d(r, _v(ax, -1.0, 0.0, 0.0)) == INF
    r = _v(ax, 8.0, 3.0, 1.0)
    assert d(r, _v(ax, 1.0, 0.0, 0.0)) == pytest.approx(3.0, rel=1e-7)
    assert d(r, _v(ax, 0.0, -1.0, 0.0)) == pytest.approx(9.0, rel=1e-7)
    assert d(r, _v(ax, -1.0, 0.0, 0.0)) == pytest.approx(5.0, rel=1e-7)
    assert d(r, unit(_v(ax, 1.0, -1.0, 0.0))) == pytest.approx(math.sqrt(2.0) * 3.0, rel=1e-7)
    r = _v(ax, 8.0, 0.5, 1.0)
    assert d(r, _v(ax, -1.0, 0.0, 0.0)) == pytest.approx(6.0, rel=1e-7)
    r = _v(ax, 10.0, 13.0, 1.0)
    assert d(r, _v(ax, -1.0, 0.0, 0.0)) == INF
    assert d(r, _v(ax, 1.0, 0.0, 0.0)) == INF
    assert d(r, unit(_v(ax, 1.0, -1.5, 0.0))) == INF
    assert d(r, unit(_v(ax, 1.0, -1.0, 0.0))) == INF
    r = _v(ax, 16.0, 4.0, 1.0)
    assert d(r, unit(_v(ax, -5.0, -2.0, 0.0))) == pytest.approx(math.sqrt(29.0), rel=1e-7)


def test_trunc_cone_rejected_definitions():
    base = dict(id=3, type="zTruncCone", vertex=[0.0, 0.0, 0.0], angle=30.0, hMin=1.0, hMax=2.0)
    for bad, msg in [(dict(angle=90.0), "Opening angle"), (dict(hMin=2.0), "hMin is greater"),
                     (dict(hMin=-1.0), "different signs"), (dict(vertex=[0.0, 0.0]), "Vertex")]:
        with pytest.raises(ValueError, match=msg):
            surf(**{**base, **bad})


def test_trunc_cone_distance_closed_form():
    """Rays from inside the truncated cone against the exit point found independently: the first
    sign change of evaluate along the ray, refined by bisection."""
    S, _ = surf(type="zTruncCone", vertex=[0.3, -0.2, 0.1], angle=35.0, hMin=1.0, hMax=4.0)
    rng = np.random.default_rng(11)
    n_hit = 0
    while n_hit < 120:
        r = list(rng.uniform(-3.0, 4.5, 3))
        if O.surf_evaluate(S, r) > -1e-3:
            continue
        u = unit(rng.normal(size=3))
        d = O.surf_distance(S, r, u)
        f = lambda t: O.surf_evaluate(S, [r[i] + t * u[i] for i in range(3)])
        ts = np.linspace(1e-9, 12.0, 4001)
        vals = np.array([f(t) for t in ts])
        sg = np.sign(vals)
        ch = np.nonzero(sg[1:] != sg[:-1])[0]
        if len(ch) == 0:
            assert d == INF, (r, u, d)
            continue
        lo, hi = ts[ch[0]], ts[ch[0] + 1]
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            if np.sign(f(mid)) == sg[ch[0]]:
                lo = mid
            else:
                hi = mid
        assert d == pytest.approx(lo, abs=1e-6), (r, u)
        n_hit += 1


def test_trunc_cone_ray_missing_the_cone_as_written():
    """truncCone distance treats a ray that misses both nappes (delta < 0, 'should never happen')
    as inside the cone laterally, so from outside it returns the entry into the axial slab: a ray
    above the top cap passing wide of the cone gets the distance to the hMax plane."""
    S, _ = surf(type="zTruncCone", vertex=[0.0, 0.0, 0.0], angle=35.0, hMin=1.0, hMax=4.0)
    u = unit([1.0, 0.0, -0.05])
    r = [-10.0, 8.0, 4.5]
    t_plane = (4.0 - 4.5) / u[2]
    assert O.surf_distance(S, r, u) == pytest.approx(t_plane, rel=1e-12)


# ---- wedge_test.f90 -------------------------------------------------------------------------------
def _wedge(ax):
    o = _v(ax, 5.0, 1.0, 0.0)
    return surf(id=8, type="xyz"[ax] + "Wedge", origin=o, halfwidth=4.0, altitude=3.0, opening=30.0,
                rotation=60.0)


@pytest.mark.parametrize("ax", AXES)
def test_wedge(ax):
    S, bb = _wedge(ax)
    ref = _v(ax, 1.0, 1.0, 0.0) + _v(ax, 9.0, 4.0, 6.0 / math.sqrt(3.0))
    assert bb == pytest.approx(ref, abs=1e-6)
    assert S.tol == pytest.approx(SURF_TOL * 4.0, abs=1e-6)
    tol = S.tol
    hs = lambda r, u: O.surf_halfspace(S, r, u)
    # testHalfspace
    assert hs(_v(ax, 7.0, 3.0, 6.0), _v(ax, 0.0, 1.0, 0.0))
    r = _v(ax, 7.0, 3.0, math.sqrt(3.0))
    assert not hs(r, _v(ax, 0.0, 1.0, 0.0))
    r = _v(ax, 7.0, 1.0, math.sqrt(3.0))
    u = unit(_v(ax, 2.0, 1.0, 0.0))
    assert not hs(r, u)
    u = unit(_v(ax, 2.0, -1.0, 0.0))
    assert hs(r, u)
    assert hs(_mv(r, 0.01 * tol, u), u)
    assert not hs(_mv(r, 0.01 * tol, u), [-x for x in u])
    assert hs(_mv(r, 2.0 * tol, u), u)
    assert hs(_mv(r, 0.5, u), u)
    assert not hs(_mv(r, -0.5, u), u)
    assert not hs(_mv(r, -0.01 * tol, u), [-x for x in u])
    u2 = _v(ax, 0.0, 0.0, -1.0)
    assert hs(_mv(r, 0.01 * tol, u2), u2)
    r = _v(ax, 9.0, 2.0, 2.0)
    u = unit(_v(ax, 1.0, -1.0, -1.0))
    assert hs(r, u)
    assert hs(_mv(r, -0.5 * tol, u), u)
    assert not hs(_mv(r, -2.0 * tol, u), u)
    assert not hs(_mv(r, -0.5, u), u)
    assert hs(_mv(r, 0.5, u), u)
    assert hs(r, _v(ax, 0.0, 0.0, -1.0))
    # testDistance
    d = lambda r, u: O.surf_distance(S, r, u)
    u = _v(ax, 1.0 / SQRT2, 1.0 / SQRT2, 0.0)
    assert d(_v(ax, 4.0, 0.0, 2.0), u) == pytest.approx(SQRT2, rel=1e-7)
    r = _v(ax, -1.0, 0.0, 2.0)
    assert d(r, u) == pytest.approx(2.0 * SQRT2, rel=1e-7)
    assert d(r, _v(ax, 1.0, 0.0, 0.0)) == INF
    assert d(r, unit(_v(ax, 1.0, -1.0, -1.0))) == INF
    assert d(r, unit(_v(ax, 0.0, 3.0, 4.0))) == INF
    s3 = math.sqrt(3.0)
    r = _v(ax, 4.0, 0.5 * (7.0 + s3), 0.5 * (5.0 * s3 - 1.0))
    assert d(r, unit(_v(ax, 0.0, -0.5, -0.5 * s3))) == pytest.approx(2.0, rel=1e-12)
    r = _v(ax, 4.0, 1.0, 3.0 * 0.5 / s3)
    assert d(r, _v(ax, 0.0, 1.0, 0.0)) == pytest.approx(1.5, rel=1e-7)
    assert d(r, _v(ax, 0.0, -1.0, 0.0)) == INF
    assert d(r, _v(ax, 0.0, 0.0, 1.0)) == INF
    r = _v(ax, 4.0, 1.0 - 0.1 * tol, 3.0 * 0.5 / s3)
    assert d(r, _v(ax, 0.0, 1.0, 0.0)) == pytest.approx(1.5 + 0.1 * tol, rel=1e-7)
    r = _v(ax, 4.0, 1.0 + 0.1 * tol, 3.0 * 0.5 / s3)
    assert d(r, _v(ax, -1.0, 0.0, 0.0)) == INF
    r = _v(ax, 7.0, 3.0, s3)
    assert d(r, _v(ax, -1.0, 0.0, 0.0)) == pytest.approx(6.0, rel=1e-7)
    assert d(r, _v(ax, 0.0, -1.0, 0.0)) == pytest.approx(2.0, rel=1e-7)
    assert d(r, unit(_v(ax, -3.0, 0.5, 0.0))) == pytest.approx(math.sqrt(37.0), rel=1e-7)
    r = _v(ax, 10.0, 6.0, 1.0)
    assert d(r, _v(ax, -1.0, 0.0, 0.0)) == INF
    assert d(r, unit(_v(ax, -1.0, -0.1, 0.0))) == INF
    assert d(r, unit(_v(ax, 0.0, -1.0, 1.0))) == INF


def test_wedge_rejected_definitions_and_root_bcs():
    base = dict(id=3, type="zWedge", origin=[0.0, 0.0, 0.0], halfwidth=1.0, altitude=1.0, opening=30.0)
    for bad, msg in [(dict(opening=0.0), "Opening angle"), (dict(rotation=400.0), "Rotation angle"),
                     (dict(halfwidth=-1.0), "halfwidth"), (dict(altitude=0.0), "altitude")]:
        with pytest.raises(ValueError, match=msg):
            surf(**{**base, **bad})


# ---- rotationMatrix (genericProcedures.f90:1098-1140) ------------------------------------------
@pytest.mark.parametrize("angles", [(90.0, 90.0, 90.0), (30.0, 45.0, 60.0), (0.0, 180.0, 359.0), (12.5, 0.0, 0.0),
                                    (200.0, 30.0, 200.0)])
def test_rotation_matrix_is_zxz_euler_as_written(angles):
    """The matrix maps global to local coordinates: the transpose of the z-x-z intrinsic rotation,
    except that the reference writes element (1,3) as sin(psi) sin(phi) where the rotation has
    sin(psi) sin(theta); the restatement keeps the reference's element, so the matrix is a
    rotation exactly when sin(phi) sin(psi) = sin(theta) sin(psi)."""
    phi, theta, psi = (math.radians(a) for a in angles)
    R = np.array(T.rotation_matrix(*angles)).reshape(3, 3)
    ref = Rotation.from_euler("ZXZ", angles, degrees=True).as_matrix().T
    mask = np.ones((3, 3), dtype=bool)
    mask[0, 2] = False
    assert np.allclose(R[mask], ref[mask], atol=1e-15)
    assert R[0, 2] == pytest.approx(math.sin(psi) * math.sin(phi), abs=1e-15)
    if abs(math.sin(phi) - math.sin(theta)) < 1e-15 or abs(math.sin(psi)) < 1e-15:
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-15)
    else:
        assert not np.allclose(R @ R.T, np.eye(3), atol=1e-6)
    with pytest.raises(ValueError):
        T.rotation_matrix(360.0, 0.0, 0.0)


# ---- simpleCell_test.f90 (one cell in a cellUniverse) -------------------------------------------
def _single_cell():
    surfs = {"b": {"id": 500, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [10.0, 10.0, 10.0]},
             "surf1": {"id": 13, "type": "sphere", "origin": [0.0, 0.0, 0.0], "radius": 2.0},
             "surf2": {"id": 4, "type": "xPlane", "x0": 0.0},
             "surf3": {"id": 99, "type": "yPlane", "y0": 0.0}}
    cells = {"c": {"id": 2, "type": "simpleCell", "surfaces": [-13, 4, 99], "filltype": "outside"}}
    unis = {"root": {"id": 1, "type": "rootUniverse", "border": 500, "fill": "u<7>"},
            "cu": {"id": 7, "type": "cellUniverse", "cells": [2]}}
    return geometry_problem(unis, surfs, cells=cells)


def test_simple_cell_inside_and_distance():
    comp = _single_cell()
    u = [1.0, 0.0, 0.0]
    assert O.uni_findcell(comp, 2, [1.0, 0.3, 0.1], u) == 1
    assert O.uni_findcell(comp, 2, [0.3, 1.3, 0.4], u) == 1
    assert O.uni_findcell(comp, 2, [-0.1, 1.2, 0.1], u) == 2      # UNDEF local cell
    assert O.uni_findcell(comp, 2, [1.8, 0.8, 0.8], u) == 2
    # surfaces are tabled in order of first reference: 13 -> 1, 4 -> 2, 99 -> 3
    r = [0.3, 0.4, 0.0]
    d, s = O.uni_distance(comp, 2, 1, r, [-1.0, 0.0, 0.0])
    assert d == pytest.approx(0.3, rel=1e-7) and s == 2
    d, s = O.uni_distance(comp, 2, 1, r, [0.0, -1.0, 0.0])
    assert d == pytest.approx(0.4, rel=1e-7) and s == 3
    d, s = O.uni_distance(comp, 2, 1, r, [1.0, 0.0, 0.0])
    assert d == pytest.approx(1.659591794, rel=1e-7) and s == 1


# ---- unionCell_test.f90 (one cell in a cellUniverse) --------------------------------------------
UNION_SURFS = {"b": {"id": 500, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [10.0, 10.0, 10.0]},
               "surf1": {"id": 13, "type": "sphere", "origin": [0.0, 0.0, 0.0], "radius": 5.0},
               "surf2": {"id": 15, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [4.6, 4.6, 4.6]},
               "surf3": {"id": 4, "type": "xTruncCylinder", "origin": [0, 0, 0], "halfwidth": 5.1, "radius": 1.0},
               "surf4": {"id": 5, "type": "yTruncCylinder", "origin": [0, 0, 0], "halfwidth": 5.1, "radius": 1.0},
               "surf5": {"id": 6, "type": "zTruncCylinder", "origin": [0, 0, 0], "halfwidth": 5.1, "radius": 1.0}}


def _union_cell(definition):
    cells = {"c": {"id": 3, "type": "unionCell", "surfaces": definition, "filltype": "outside"}}
    unis = {"root": {"id": 1, "type": "rootUniverse", "border": 500, "fill": "u<7>"},
            "cu": {"id": 7, "type": "cellUniverse", "cells": [3]}}
    return geometry_problem(unis, UNION_SURFS, cells=cells)


EASY = ["-13", "-15"]
COMPLEX = ["<", "-13", "-15", ">", "#", "<", "-4", ":", "-5", ":", "-6", ">"]


def test_union_cell_inside():
    easy, cplx = _union_cell(EASY), _union_cell(COMPLEX)
    u = [1.0, 0.0, 0.0]
    inside = lambda comp, r: O.uni_findcell(comp, 2, r, u) == 1
    assert inside(easy, [-4.59, 0.0, 0.0])
    assert inside(easy, [0.3, 1.3, 0.4])
    assert inside(cplx, [3.0, 2.3, 0.4])
    assert inside(cplx, [-1.1, 1.05, 0.0])
    assert not inside(easy, [0.0, 4.9, 0.0])
    assert not inside(easy, [4.58, 4.58, -4.58])
    assert not inside(cplx, [0.1, 0.3, -0.2])
    assert not inside(cplx, [-4.59, 0.0, 0.25])
    assert not inside(cplx, [4.58, 4.58, -4.58])


def test_union_cell_distance():
    easy, cplx = _union_cell(EASY), _union_cell(COMPLEX)
    # surfaces are tabled in order of first reference: easy 13 -> 1, 15 -> 2 (the root box is
    # a box-table entry); complex 13, 15, 4, 5, 6 -> 1..5
    d, s = O.uni_distance(easy, 2, 1, [0.2, -0.1, 0.0], [-0.980580676, 0.0, 0.196116135])
    assert d == pytest.approx(4.895058733, rel=1e-6) and s == 2
    d, s = O.uni_distance(easy, 2, 1, [6.0, 0.0, 0.0], [-1.0, 0.0, 0.0])
    assert d == pytest.approx(1.0, rel=1e-6) and s == 1
    d, s = O.uni_distance(easy, 2, 1, [4.5, 4.5, 4.5], [-0.577350269] * 3)
    assert d == pytest.approx(2.794228634, rel=1e-6) and s == 1
    d, s = O.uni_distance(cplx, 2, 1, [1.2, 0.0, 0.0], [0.0, -1.0, 0.0])
    assert d == pytest.approx(1.0, rel=1e-6) and s == 3


@pytest.mark.parametrize("definition,msg", [
    (["1", ":", "2", "3"], "ambiguous"), (["<", "1", ":", "2"], "parentheses"),
    (["1", ":"], "ends with a union"), (["1", "#"], "ends with a complement"),
    (["1", "?"], "Unrecognised"), (["1", "-1"], "repeated surfaces"),
])
def test_union_cell_rejected_definitions(definition, msg):
    with pytest.raises(ValueError, match=msg):
        T.union_cell_tokens(definition, 9, lambda sid: sid - 1)


def test_union_cell_de_morgan_tokens():
    U, I, OP, CL = T.U_UNION, T.U_INTERSECT, T.U_OPEN, T.U_CLOSE
    # #<1 #<2 3>> = not 1 or (2 and 3)
    simple, tok = T.union_cell_tokens(["#", "<", "1", "#", "<", "2", "3", ">", ">"], 1, lambda sid: sid - 1)
    assert simple == 0 and tok == [OP, -1, U, OP, 2, I, 3, CL, CL]
    simple, tok = T.union_cell_tokens(["-1", "#", "2"], 1, lambda sid: sid - 1)
    assert simple == 1 and tok == [-1, I, -2]


def test_union_cell_matches_the_boolean_formula():
    """Random expressions over six planes: the cell's inside equals the formula evaluated with each
    surface's halfspace (complements, brackets, short circuits and deMorgan all exercised)."""
    rng = np.random.default_rng(5)
    planes = {f"p{k}": {"id": 20 + k, "type": "plane", "coeffs": list(rng.normal(size=3)) + [rng.normal() * 0.3]}
              for k in range(6)}

    def expr(depth):
        n = int(rng.integers(1, 4))
        op = ":" if rng.random() < 0.5 else ""
        items, py = [], []
        for k in range(n):
            neg = rng.random() < 0.3
            if depth > 0 and rng.random() < 0.4:
                sub, subpy = expr(depth - 1)
                items += (["#"] if neg else []) + ["<"] + sub + [">"]
                py.append(("not " if neg else "") + "(" + subpy + ")")
            else:
                sid = int(rng.integers(0, 6))
                sign = -1 if rng.random() < 0.5 else 1
                if neg:
                    items += ["#", str(sign * (20 + sid))]
                else:
                    items.append(str(sign * (20 + sid)))
                term = f"({'' if sign > 0 else 'not '}h[{sid}])"
                py.append(("not " if neg else "") + term)
            if k < n - 1 and op:
                items.append(op)
        return items, (" or " if op else " and ").join(py)

    n_checked = 0
    for trial in range(40):
        items, py = expr(2)
        surfs = dict(planes)
        surfs["b"] = {"id": 500, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [3.0, 3.0, 3.0]}
        cells = {"c": {"id": 3, "type": "unionCell", "surfaces": items, "filltype": "outside"}}
        unis = {"root": {"id": 1, "type": "rootUniverse", "border": 500, "fill": "u<7>"},
                "cu": {"id": 7, "type": "cellUniverse", "cells": [3]}}
        try:
            comp = geometry_problem(unis, surfs, cells=cells)
        except ValueError as e:          # an intersection-only cell naming a surface twice
            assert "repeated surfaces" in str(e)
            continue
        S = [T.compile_surface({**planes[f"p{k}"]}, [])[0] for k in range(6)]
        for _ in range(25):
            r = list(rng.uniform(-2.0, 2.0, 3))
            u = unit(rng.normal(size=3))
            h = [bool(O.surf_halfspace(S[k], r, u)) for k in range(6)]
            want = eval(py, {"h": h})
            assert (O.uni_findcell(comp, 2, r, u) == 1) == want, (items, py, r)
            n_checked += 1
    assert n_checked >= 500


# ---- cellUniverse_test.f90 --------------------------------------------------------------------
@pytest.fixture(scope="module")
def cell_uni():
    surfs = {"b": {"id": 500, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [50.0, 50.0, 50.0]},
             "surf1": {"id": 1, "type": "sphere", "origin": [0.0, 0.0, 0.0], "radius": 2.0},
             "surf2": {"id": 2, "type": "sphere", "origin": [4.0, 0.0, 0.0], "radius": 1.0}}
    cells = {"cell1": {"id": 1, "type": "simpleCell", "surfaces": [-1], "filltype": "uni", "universe": 3},
             "cell2": {"id": 2, "type": "simpleCell", "surfaces": [1, 2], "filltype": "uni", "universe": 4}}
    unis = {"root": {"id": 100, "type": "rootUniverse", "border": 500, "fill": "u<1>"},
            "u1": {"id": 1, "type": "cellUniverse", "origin": [2.0, 0.0, 0.0], "rotation": [90.0, 90.0, 90.0],
                   "cells": [1, 2]},
            "u3": {"id": 3, "type": "pinUniverse", "radii": [0.0], "fills": ["m1"]},
            "u4": {"id": 4, "type": "pinUniverse", "radii": [0.0], "fills": ["m2"]}}
    return geometry_problem(unis, surfs, cells=cells, materials=("m1", "m2"))


def test_cell_universe_fill(cell_uni):
    fill, _ = cell_uni.graph
    # root [u<1>, outside], cellUniverse [u<3>, u<4>, UNDEF, OVERLAP] (fills as uniIdx)
    assert list(fill[2:6]) == [-3, -4, T.UNDEF_MAT, T.OVERLAP_MAT]


def test_cell_universe_enter_rotated(cell_uni):
    TOL = 1e-7
    o = O.place(cell_uni, [0.0, 0.0, 3.0], [0.0, 0.0, 1.0])
    assert o.nesting == 3 and o.localID[1] == 1 and o.matIdx == 1
    assert list(o.r[1]) == pytest.approx([1.0, 0.0, 0.0], abs=TOL)
    o = O.place(cell_uni, [2.0, 0.0, 1.0], [0.0, 1.0, 0.0])
    assert o.localID[1] == 2 and o.matIdx == 2
    assert list(o.r[1]) == pytest.approx([-1.0, 0.0, 2.0], abs=TOL)
    o = O.place(cell_uni, [0.0, 0.0, 6.5], [1.0, 0.0, 0.0])
    assert o.nesting == 2 and o.localID[1] == 3 and o.matIdx == T.UNDEF_MAT
    assert list(o.r[1]) == pytest.approx([4.5, 0.0, 0.0], abs=TOL)


def test_cell_universe_distance_and_undefined(cell_uni):
    TOL = 1e-7
    d, s = O.uni_distance(cell_uni, 2, 1, [-1.0, 0.0, 0.0], [1.0, 0.0, 0.0])
    assert d == pytest.approx(3.0, rel=TOL) and s == 1
    d, s = O.uni_distance(cell_uni, 2, 2, [7.0, 0.0, 0.0], [-1.0, 0.0, 0.0])
    assert d == pytest.approx(2.0, rel=TOL) and s == 2
    d, s = O.uni_distance(cell_uni, 2, 2, [7.0, 0.0, 0.0], [1.0, 0.0, 0.0])
    assert d == INF and s == 0
    # the UNDEF (n+1) and OVERLAP (n+2) local cells have no distance: fatalError in the reference
    for local, msg in ((3, "undefined local cell"), (4, "overlapping local cell")):
        O.place(cell_uni, [0.0, 0.0, 3.0])          # clears the error state
        O.uni_distance(cell_uni, 2, local, [0.0, 0.0, 0.0], [1.0, 0.0, 0.0])
        with pytest.raises(RuntimeError, match=msg):
            O.check(O.lib().orc_error_code())


def test_cell_universe_cross(cell_uni):
    # crossing surface 1 from cell 1 at (0, 2, 0) moving +y: NUDGE, then the search finds cell 2;
    # the ST move from inside cell 1 reaches the same crossing through geometryStd%move
    o = O.place(cell_uni, [0.0, 0.0, 2.0], [0.0, 1.0, 0.0])      # local (0, 0, 0) in cell 1
    assert o.localID[1] == 1 and list(o.r[1]) == pytest.approx([0.0, 0.0, 0.0], abs=1e-15)
    o, u, d, ev = O.move(cell_uni, [0.0, 0.0, 2.0], [0.0, 1.0, 0.0], 10.0)
    assert ev == 3 and o.localID[1] == 2 and o.matIdx == 2
    assert d == pytest.approx(2.0, rel=1e-12)
    # the local position carries the nudge of 1e-8 along the local direction (-y after rotation)
    assert o.r[1][1] == pytest.approx(-2.0 - 1e-8, abs=1e-14)


def test_cell_universe_overlap():
    surfs = {"b": {"id": 500, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [50.0, 50.0, 50.0]},
             "surf1": {"id": 1, "type": "sphere", "origin": [21.0, 0.0, 0.0], "radius": 4.5},
             "surf2": {"id": 2, "type": "sphere", "origin": [0.0, 0.0, 0.0], "radius": 24.1}}
    cells = {"cell1": {"id": 1, "type": "simpleCell", "surfaces": [-1], "filltype": "mat", "material": "m1"},
             "cell2": {"id": 2, "type": "simpleCell", "surfaces": [1, -2], "filltype": "mat", "material": "m1"},
             "cell3": {"id": 3, "type": "simpleCell", "surfaces": [2], "filltype": "mat", "material": "m1"}}
    unis = {"root": {"id": 100, "type": "rootUniverse", "border": 500, "fill": "u<2>"},
            "u2": {"id": 2, "type": "cellUniverse", "checkOverlap": 1, "cells": [1, 2, 3]}}
    comp = geometry_problem(unis, surfs, cells=cells)
    u = [1.0, 0.0, 0.0]
    assert O.uni_findcell(comp, 2, [22.0, 0.0, 0.0], u) == 1
    assert O.uni_findcell(comp, 2, [0.0, 25.0, 0.0], u) == 3
    assert O.uni_findcell(comp, 2, [24.2, 0.0, 0.0], u) == 5      # in cells 1 and 3: OVERLAP
    # without checkOverlap the first cell in definition order wins
    unis["u2"]["checkOverlap"] = 0
    comp = geometry_problem(unis, surfs, cells=cells)
    assert O.uni_findcell(comp, 2, [24.2, 0.0, 0.0], u) == 1


def test_cell_definitions_rejected():
    surfs = {"b": {"id": 500, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [5.0, 5.0, 5.0]},
             "s": {"id": 1, "type": "sphere", "origin": [0.0, 0.0, 0.0], "radius": 1.0}}
    unis = {"root": {"id": 100, "type": "rootUniverse", "border": 500, "fill": "u<2>"},
            "u2": {"id": 2, "type": "cellUniverse", "cells": [1]}}
    bad = {"c": {"id": 1, "type": "simpleCell", "surfaces": [1, -1], "filltype": "mat", "material": "m1"}}
    with pytest.raises(ValueError, match="repeated surfaces"):
        geometry_problem(unis, surfs, cells=bad)
    bad = {"c": {"id": 1, "type": "simpleCell", "surfaces": [-1], "filltype": "pie"}}
    with pytest.raises(ValueError, match="Unknown type of cell filling"):
        geometry_problem(unis, surfs, cells=bad)
    bad = {"c": {"id": 1, "type": "fancyCell", "surfaces": [-1], "filltype": "outside"}}
    with pytest.raises(NotImplementedError):
        geometry_problem(unis, surfs, cells=bad)


# ---- a sphere as the root border (vacuum), closed form ------------------------------------------
def test_absorbing_sphere_escape_closed_form():
    """Pure absorber sphere of radius R, centred isotropic point source: every history either
    collides once (probability 1 - exp(-Sigma R)) or leaks, so the collision-estimator flux per
    source neutron is (1 - exp(-Sigma R)) / Sigma."""
    R, sig = 1.7, 0.9
    pop, cycles = 4000, 20
    for top in ("transportOperatorST", "transportOperatorDT"):
        inp, mg = configs.absorbing_sphere(R, sig, pop, cycles, top)
        comp = T.compile_problem(inp, mg)
        src = T.compile_source(inp["source"], comp)
        res, done, h = O.fixed_run(comp, src, pop, cycles, 3, hist=True, threads=4)
        expect = pop * (1.0 - math.exp(-sig * R)) / sig
        z = (res[0, 0] - expect) / res[0, 1]
        assert abs(z) < 3, (top, res[0], expect)
        assert set(np.unique(h["fate"])) <= {5001, 5002}


# ---- trackClerk (trackClerk_class.f90 reportPath) -------------------------------------------------
def test_track_clerk_path_lengths_replayed():
    """In the absorbing sphere each history's only path is min(d, R) with d = -ln(u)/Sigma and u
    the first number of the history's stream (source sampling uses its own stream), so the
    track-length tally of a batch is the sum of those paths, replayed here with Python floats."""
    R, sig, pop, cycles, seed = 1.7, 0.9, 500, 2, 3
    inp, mg = configs.absorbing_sphere(R, sig, pop, cycles, "transportOperatorST")
    comp = T.compile_problem(inp, mg)
    assert [c["kind"] for c in comp.clerk_info] == ["collisionClerk", "trackClerk"]
    src = T.compile_source(inp["source"], comp)
    res, done, _ = O.fixed_run(comp, src, pop, cycles, seed, threads=2)
    per_batch = []
    for b in range(cycles):
        tot = 0.0
        for j in range(1, pop + 1):
            u, _ = O.rng_get(O.rng_skip(seed, 152917 * (2 * pop * b + pop + j)))
            tot += min(-math.log(u) / sig, R)
        per_batch.append(tot)
    assert res[1, 0] == pytest.approx(sum(per_batch) / cycles, rel=1e-12)
    # and both estimators estimate (1 - exp(-sig R)) / sig per source neutron
    expect = pop * (1.0 - math.exp(-sig * R)) / sig
    for k in (0, 1):
        assert abs(res[k, 0] - expect) < 4 * res[k, 1] + 1e-9 * expect


def test_track_clerk_scores_nothing_under_delta_tracking():
    inp, mg = configs.absorbing_sphere(1.7, 0.9, 300, 2, "transportOperatorDT")
    comp = T.compile_problem(inp, mg)
    src = T.compile_source(inp["source"], comp)
    res, _, _ = O.fixed_run(comp, src, 300, 2, 3, threads=2)
    assert res[0, 0] > 0 and res[1, 0] == 0.0


def test_track_and_collision_estimators_agree_in_a_slab():
    """Region-integrated fluxes: a path never crosses a region boundary (surface tracking stops
    there), so the track-length estimator of a material's flux is unbiased like the collision
    estimator (a finer mesh is not: trackClerk puts a whole path in its starting bin)."""
    pop, cycles = 4000, 10
    inp, mg = configs.slab_fixed(pop=pop, cycles=cycles, transport="transportOperatorST", track=True)
    inp["tally"]["track_total"] = {"type": "trackClerk", "response": ["flux"], "flux": {"type": "fluxResponse"}}
    flt = inp["tally"]["track_mats"]["filter"]
    inp["tally"]["coll_mats"] = dict(inp["tally"]["track_mats"], type="collisionClerk", filter=flt)
    comp = T.compile_problem(inp, mg)
    src = T.compile_source(inp["source"], comp)
    res, _, _ = O.fixed_run(comp, src, pop, cycles, inp["seed"], threads=4)
    info = {c["name"]: c for c in comp.clerk_info}
    g = lambda n: res[info[n]["addr"]: info[n]["addr"] + info[n]["size"]]
    for a, b in ((g("total"), g("track_total")), (g("coll_mats"), g("track_mats"))):
        z = (a[:, 0] - b[:, 0]) / np.hypot(a[:, 1], b[:, 1])
        assert np.all(np.abs(z) < 4), z
    # the mesh sums to the total flux for both estimators
    assert g("mesh")[:, 0].sum() == pytest.approx(g("total")[0, 0], rel=1e-9)
    assert g("track_mesh")[:, 0].sum() == pytest.approx(g("track_total")[0, 0], rel=1e-9)


# ---- cellMap and energyMap (Tallies/TallyMaps/Maps1D/{cellMap,energyMap}_class.f90) ----------------
def test_cell_map_sums_to_the_material_map():
    """Collision and track estimators binned by cell add up, cell by cell, to the same estimators
    binned by each cell's material; a multigroup energyMap bins nothing."""
    inp, mg = configs.cell_lattice(pop=3000, transport="transportOperatorST", shapes=True)
    comp = T.compile_problem(inp, mg)
    bank = O.fission_source(comp, 3000, 4242)
    _, bins, _, _ = O.eigen_cycle(comp, bank, 777, 0.95)
    info = {c["name"]: c for c in comp.clerk_info}
    cells = inp["geometry"]["cells"]
    ids = [int(cd["id"]) for cd in cells.values()]
    mat_of = {int(cd["id"]): cd.get("material") for cd in cells.values()}
    order = ["fuel", "structural", "moderator"]
    for cell_clerk, mat_clerk, width, col in (("cells", "mats", 1, 0), ("track_cells", "track_mats", 2, 0)):
        cb = bins[info[cell_clerk]["addr"]: info[cell_clerk]["addr"] + info[cell_clerk]["size"]]
        mb = bins[info[mat_clerk]["addr"]: info[mat_clerk]["addr"] + info[mat_clerk]["size"]]
        assert cb[-1] == 0.0                                     # every point is in a defined cell
        for m, name in enumerate(order):
            s = sum(cb[i] for i, cid in enumerate(ids) if mat_of[cid] == name)
            assert s == pytest.approx(mb[m * width + col], rel=1e-12), (cell_clerk, name)
        assert (cb[:-1] > 0).sum() >= 10
    e = info["energy"]
    assert e["size"] == 4 and not bins[e["addr"]: e["addr"] + 4].any()


def test_energy_and_cell_map_prints():
    from scone_amd import output as OUT
    inp, mg = configs.cell_lattice(pop=100, shapes=True)
    comp = T.compile_problem(inp, mg)
    out = OUT.OutputFile()
    OUT.print_tally(out, comp, np.zeros((int(comp.problem.n_bins), 2)))
    text = out.text()
    line = [l for l in text.splitlines() if l.startswith("energy_EnergyBounds")][0]
    vals = [float(x) for x in line.split("[ ")[1].split("]")[0].split(",")]
    edges = [1e-11]
    for _ in range(4):
        edges.append(edges[-1] * math.exp(math.log(20.0 / 1e-11) / 4))
    assert line.endswith("],4,2);")
    assert vals[:4] == pytest.approx(edges[:4], rel=1e-5) and vals[4:] == pytest.approx(edges[1:], rel=1e-5)
    line = [l for l in text.splitlines() if l.startswith("cells_CellBins")][0]
    assert line.startswith("cells_CellBins = reshape([ {'1'},{'2'},") and "{'undefined'}],1," in line


def test_map_definitions_rejected():
    inp, mg = configs.cell_lattice(pop=100, shapes=True)
    inp["activeTally"]["energy"]["map"] = {"type": "energyMap", "grid": "predef", "name": "wims69"}
    with pytest.raises(NotImplementedError):
        T.compile_problem(inp, mg)
    inp, mg = configs.cell_lattice(pop=100, shapes=True)
    inp["activeTally"]["cells"]["map"]["cells"] = [1, 999]
    with pytest.raises(ValueError, match="Cell ID 999"):
        T.compile_problem(inp, mg)
