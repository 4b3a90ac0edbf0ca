"""RNG and shared-math pins (CPU).

* The LCG step and Brown's skip against Python integer arithmetic and the reference's own
  pow_of_gsq table (RandomNumbers/RNG_class.f90:61-123, fixture tests/golden/rng_pow_of_gsq.json).
* RNG_test.f90:36-70: skipping +/-13456757 reproduces the sequentially drawn numbers.
* scn_log / scn_sincos / scn_norm2 / scn_rotate against libm (numpy) and closed forms.
"""
import json
import math
import os

import numpy as np
import pytest

import binding as O

G = 2806196910506780709
MASK = (1 << 63) - 1
HERE = os.path.dirname(os.path.abspath(__file__))


def py_get(s):
    s = (G * s) & MASK
    s = (s + 1) & MASK
    return float(s) * 2.0 ** -63, s


def test_get_matches_integer_recurrence():
    s = 0x5C3A84C9
    for _ in range(2000):
        u_ref, s_ref = py_get(s)
        u, s2 = O.rng_get(s)
        assert s2 == s_ref and u == u_ref
        assert 0.0 <= u <= 1.0
        s = s2


def test_pow_of_gsq_table_is_repeated_squaring():
    fx = json.load(open(os.path.join(HERE, "golden", "rng_pow_of_gsq.json")))
    g = fx["g"]
    vals = fx["pow_of_gsq"]
    x = g
    for i in range(63):
        x = (x * x) & MASK
        assert x == vals[i], i
    # skip(2**i) from state s is G**(2**i) * s + C: check against the table for the first terms
    for i in range(1, 20):
        s = 12345
        direct = s
        for _ in range(2 ** min(i, 10)):
            direct = py_get(direct)[1]
        assert O.rng_skip(s, 2 ** min(i, 10)) == direct


@pytest.mark.parametrize("k", [0, 1, 2, 3, 17, 152917, 999983, 2 ** 40 + 7])
def test_skip_closed_form(k):
    s = 0x5C3A84C9
    # S_k = g^k s + (g^k - 1)/(g - 1) mod 2^63, evaluated with Python big integers
    M = 1 << 63
    gk = pow(G, k, M * (G - 1))
    ck = ((gk - 1) // (G - 1)) % M
    ref = (pow(G, k, M) * s + ck) % M
    assert O.rng_skip(s, k) == ref


def test_skip_forward_backward_RNG_test():
    # RandomNumbers/Tests/RNG_test.f90:36-70
    seed = 0x5C3A84C9
    r_start, s = O.rng_get(seed)
    st = s
    N = 13456757
    s2 = O.rng_skip(seed, N)
    r2_end, s2b = O.rng_get(s2)
    # rand1 draws N more numbers after the first (N + 1 steps in all): big-integer closed form
    M = 1 << 63
    k = N + 1
    x = (pow(G, k, M) * seed + ((pow(G, k, M * (G - 1)) - 1) // (G - 1))) % M
    r_end = float(x) * 2.0 ** -63
    assert r_end == r2_end
    back = O.rng_skip(s2b, -(N + 1))
    r2_start, _ = O.rng_get(back)
    assert r2_start == r_start


def ulp_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_log_within_one_ulp():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.random(20000), rng.random(2000) * 1e-300, 1 - rng.random(2000) * 1e-10,
                         rng.uniform(0.5, 2.0, 2000), [1.0, 0.5, 2.0 ** -63, 1 - 2.0 ** -53]])
    ours = np.array([O.log(float(x)) for x in xs])
    ref = np.log(xs)
    err = ulp_err(ours[ref != 0], ref[ref != 0])
    assert err.max() <= 1.0, err.max()
    assert O.log(1.0) == 0.0
    assert O.log(0.0) == -math.inf


def test_sincos_within_two_ulp_on_the_azimuth_range():
    rng = np.random.default_rng(2)
    xs = 2.0 * math.pi * rng.random(20000)
    s = np.empty_like(xs)
    c = np.empty_like(xs)
    for i, x in enumerate(xs):
        s[i], c[i] = O.sincos(float(x))
    # absolute error relative to 1 ulp of the magnitude-1 range is what the rotation sees
    assert np.max(np.abs(s - np.sin(xs))) < 4e-16
    assert np.max(np.abs(c - np.cos(xs))) < 4e-16
    assert O.sincos(0.0) == (0.0, 1.0)


def test_norm2_and_rotation_properties():
    rng = np.random.default_rng(3)
    for _ in range(500):
        v = rng.normal(size=3)
        assert abs(O.norm2(v) - float(np.linalg.norm(v))) <= 2e-16 * np.linalg.norm(v) * 4
        d = v / np.linalg.norm(v)
        mu = rng.uniform(-1, 1)
        phi = rng.uniform(0, 2 * math.pi)
        n = np.array(O.rotate(d, mu, phi))
        assert abs(np.linalg.norm(n) - 1.0) < 1e-15
        assert abs(float(n @ d) - mu) < 1e-12      # polar cosine preserved
    # pole branch (|w| ~ 1): rotateVector switches to the v-based frame (genericProcedures.f90:1074)
    n = np.array(O.rotate([0.0, 0.0, 1.0], 0.3, 1.1))
    assert abs(n[2] - 0.3) < 1e-15 and abs(np.linalg.norm(n) - 1) < 1e-15


def test_canonical_sum_order():
    x = np.random.default_rng(4).random(1000)
    # chunks of 32 left to right, then pairwise
    chunks = [sum_left(x[i:i + 32]) for i in range(0, len(x), 32)]
    while len(chunks) > 1:
        chunks = [chunks[i] + chunks[i + 1] if i + 1 < len(chunks) else chunks[i] for i in range(0, len(chunks), 2)]
    assert O.dsum(x) == chunks[0]
    assert O.dsum(np.zeros(0)) == 0.0


def sum_left(v):
    s = 0.0
    for t in v:
        s = s + float(t)
    return s
