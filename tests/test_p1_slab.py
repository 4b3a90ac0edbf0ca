"""P1 scattering (multiScatterP1MG with sampleLegendre_P1) and the reference's SCONE_Slab input
(InputFiles/SCONE_Slab with XS/URRa_2_1_XSS, data fixtures under tests/golden/scone_slab), CPU.

* sampleLegendre_P1 (legendrePoly_func.f90:35-99): the sampled cosine has density
  (1 + c mu) / 2 on [-1, 1] for |c| <= 1 (mean c / 3; CDF (mu + 1)/2 + c (mu^2 - 1)/4), and for
  1 < c <= 3 an atom at mu = 1 of weight (c - 1)/2 on top of the linear density; mean c / 3
  throughout.  The draw order is replayed with the LCG in Python integers.
* URR-2-1-SL of the analytical criticality benchmark set (Sood, Forster, Parsons, LA-13511): a
  two-group slab with linearly anisotropic scattering whose half-thickness 9.4959 cm makes it
  critical for the exact linear kernel.  The deck's coefficients 3 P1/P0 are 1.317 and 1.023:
  above 1 the linear kernel is negative at backward angles, and sampleLegendre_P1 keeps the mean
  cosine c/3 with a non-negative mixture (an atom at mu = 1) instead, so the sampled transport
  problem is not exactly the analytic one.  Measured on the MI355X over 12 independent seeds:
  k = 1.00101 +/- 0.00011 (delta tracking) and 1.00112 +/- 0.00010 (surface tracking).  The test
  asks for |k - 1| < 0.003 + 3 sd.
"""
import math
import os

import numpy as np
import pytest
from scipy import stats

import binding as O
from scone_amd import dictio, tables as T

HERE = os.path.dirname(os.path.abspath(__file__))
SLAB = os.path.join(HERE, "golden", "scone_slab")
G = 2806196910506780709
MASK = (1 << 63) - 1


def py_get(s):
    s = (G * s + 1) & MASK
    return float(s) * 2.0 ** -63, s


def sample(c, n, s0=12345):
    out = np.empty(n)
    s = s0
    for i in range(n):
        out[i], s = O.legendre_p1(c, s)
    return out


@pytest.mark.parametrize("c", [-2.5, -0.6, 0.0, 0.3, 0.99, 1.0, 1.3166, 2.0, 3.0])
def test_legendre_p1_moments(c):
    n = 40000
    mu = sample(c, n)
    assert np.all(np.abs(mu) <= 1.0)
    sd = math.sqrt(1.0 / 3.0 / n)
    assert abs(mu.mean() - c / 3.0) < 5 * sd, (mu.mean(), c / 3.0)
    if abs(c) <= 1.0:
        cdf = lambda x: (x + 1.0) / 2.0 + c * (x * x - 1.0) / 4.0
        assert stats.kstest(mu, cdf).pvalue > 1e-4
    else:
        atom = np.mean(mu == math.copysign(1.0, c))
        w = (abs(c) - 1.0) / 2.0
        assert abs(atom - w) < 5 * math.sqrt(w * (1 - w) / n) + 1e-12


def test_legendre_p1_draw_order():
    for c, s0 in [(0.4, 99), (2.2, 1234), (-0.7, 77777)]:
        for k in range(50):
            s = s0 + 7919 * k
            mu, s_end = O.legendre_p1(c, s)
            u1, t = py_get(s)
            a = abs(c)
            thr, top, low = (a, "lin", "uni") if a < 1 else (0.5 * (a - 1.0), "delta", "lin")
            ex = top if u1 < thr else low
            if ex == "uni":
                u2, t = py_get(t)
                x = 2.0 * u2 - 1.0
            elif ex == "lin":
                u2, t = py_get(t)
                x = 2.0 * math.sqrt(u2) - 1.0
            else:
                x = 1.0
            if c < 0:
                x = -x
            assert mu == x and s_end == t


def test_legendre_p1_rejects_c_above_3():
    with pytest.raises(RuntimeError, match="P1 must have absolute value"):
        O.legendre_p1(3.5, 1)


def slab():
    inp = dictio.read(os.path.join(SLAB, "SCONE_Slab"))
    return inp, T.compile_problem(inp, base_dir=SLAB)


def test_scone_slab_compiles_with_p1_coefficients():
    inp, comp = slab()
    m = comp.mats[0]
    # XS/URRa_2_1_XSS: P0 = (0.62568 0.029227; 0 2.44383), P1 = (0.27459 0.0075737; 0 0.83318)
    assert m.P1 == [0.27459 / 0.62568 * 3.0, 0.0075737 / 0.029227 * 3.0, 0.0, 0.83318 / 2.443830 * 3.0]
    assert comp.problem.transport == T.TRANSPORT_DT
    assert comp.problem.norm_bin == 0 and comp.problem.norm_value == 100.0
    assert comp.bounds[0] == -9.4959 and comp.bounds[3] == 9.4959


def test_scone_slab_is_critical():
    inp, comp = slab()
    pop = 10000
    bank = O.fission_source(comp, pop, 3)
    R, kh, ba, _ = O.eigen_run(comp, dict(pop=pop, inactive=20, active=100, seed=3), bank, threads=8)
    k, sd = R.k_implicit
    assert abs(k - 1.0) < 0.003 + 3 * sd, (k, sd)
    assert sd < 0.002
