"""Fixed-source driver and sources of the oracle (CPU).

* pointSource / materialSource sampling (pointSource_class.f90:183-246,
  materialSource_class.f90:150-231) replayed draw by draw from independently pinned pieces: the
  per-particle stream (source_inter.f90:110-116, stride(i)), the LCG recurrence in Python
  integers, rotateVector and whatIsAt (pinned in test_rng_math / test_geometry_oracle).
* Isotropy and probG frequencies of the sampled source (closed forms).
* The stride bookkeeping of fixedSourcePhysicsPackage%cycles: one skip by 2*pop*b + j equals the
  package's sequence of strides (what the engine relies on to start any batch directly).
* Whole-run closed forms: non-multiplying infinite medium, flux per source neutron (T - S)^-1 q;
  subcritical infinite medium (k_inf = 0.6) whose fission secondaries are transported in the
  source neutron's history, (T - S - chi nuSf^T)^-1 q; pure absorber slab with a centred
  isotropic point source, flux (1 - E2(Sigma_t a)) / Sigma_t.
"""
import math

import numpy as np
import pytest
from scipy import special, stats

import binding as O
from scone_amd import configs, tables as T

G = 2806196910506780709
MASK = (1 << 63) - 1
STRIDE = 152917


def py_get(s):
    s = (G * s + 1) & MASK
    return float(s) * 2.0 ** -63, s


def compiled(inp, mg):
    comp = T.compile_problem(inp, mg)
    return comp, T.compile_source(inp["source"], comp)


def test_point_source_draw_order_and_stream():
    inp, mg = configs.inf_medium_fixed(seed=5, nG=4)
    comp, src = compiled(inp, mg)
    master = 0x1234567
    n = 300
    bank = O.source_sample(comp, src, n, master)
    pg = inp["source"]["probG"]
    tot = 0.0
    for x in pg:
        tot = tot + x
    pgn = [x / tot for x in pg]
    for i in range(1, n + 1):
        s = O.rng_skip(master, STRIDE * i)
        u1, s = py_get(s)
        u2, s = py_get(s)
        u3, s = py_get(s)
        d = O.rotate([1.0, 0.0, 0.0], 2.0 * u1 - 1.0, 2.0 * math.pi * u2)
        g = 1
        r = u3
        for g in range(1, len(pgn) + 1):
            r = r - pgn[g - 1]
            if r < 0.0:
                break
        p = bank[i - 1]
        assert list(p["dir"]) == list(d), i
        assert p["G"] == g and p["wgt"] == 1.0 and p["time"] == 0.0
        assert list(p["r"]) == [0.0, 0.0, 0.0]
        assert p["matIdx"] == -1 and p["uniqueID"] == -1 and p["cellIdx"] == -1


def test_point_source_fixed_direction_is_normalised_by_norm2():
    inp, mg = configs.inf_medium_fixed(seed=5, nG=2)
    inp["source"] = {"type": "pointSource", "r": [1.0, -2.0, 0.5], "dir": [3.0, 4.0, 12.0], "G": 2}
    comp, src = compiled(inp, mg)
    bank = O.source_sample(comp, src, 5, 99)
    n = O.norm2([3.0, 4.0, 12.0])
    assert n == 13.0
    for p in bank:
        assert list(p["dir"]) == [3.0 / 13.0, 4.0 / 13.0, 12.0 / 13.0]
        assert p["G"] == 2 and list(p["r"]) == [1.0, -2.0, 0.5]


def test_point_source_isotropy_and_group_frequencies():
    inp, mg = configs.inf_medium_fixed(seed=6, nG=5)
    comp, src = compiled(inp, mg)
    n = 200000
    bank = O.source_sample(comp, src, n, 77)
    d = bank["dir"]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-14)
    # each component uniform in [-1, 1] for an isotropic direction: mean 0, mean square 1/3
    for k in range(3):
        assert abs(d[:, k].mean()) < 4 * math.sqrt(1.0 / 3.0 / n)
        assert abs((d[:, k] ** 2).mean() - 1.0 / 3.0) < 4 * math.sqrt(4.0 / 45.0 / n)
    assert stats.kstest(d[:, 2], "uniform", args=(-1.0, 2.0)).pvalue > 1e-4
    pg = np.asarray(inp["source"]["probG"])
    pg = pg / pg.sum()
    counts = np.bincount(bank["G"], minlength=6)[1:]
    chi2 = (((counts - n * pg) ** 2) / (n * pg)).sum()
    assert stats.chi2.sf(chi2, 4) > 1e-4


def test_material_source_rejection_replay():
    inp, mg = configs.slab_fixed(seed=7, regions=3)
    comp, src = compiled(inp, mg)
    S = src.source
    master = 0xBEEF
    n = 200
    bank = O.source_sample(comp, src, n, master)
    bottom, top = list(S.bottom), list(S.top)
    for i in range(1, n + 1):
        s = O.rng_skip(master, STRIDE * i)
        while True:
            u = []
            for _ in range(3):
                x, s = py_get(s)
                u.append(x)
            r = [(top[k] - bottom[k]) * u[k] + bottom[k] for k in range(3)]
            t, s = py_get(s)
            time = S.t_low + t * (S.t_high - S.t_low)
            pl = O.place(comp, r)
            if pl.matIdx == 0 or pl.matIdx != S.mat:
                continue
            break
        m1, s = py_get(s)
        m2, s = py_get(s)
        d = O.rotate([1.0, 0.0, 0.0], 2.0 * m1 - 1.0, 2.0 * math.pi * m2)
        p = bank[i - 1]
        assert list(p["r"]) == r and p["time"] == time and list(p["dir"]) == d
        assert p["matIdx"] == S.mat and p["uniqueID"] == pl.uniqueID and p["G"] == S.G


def test_material_source_fails_after_200_attempts():
    inp, mg = configs.slab_fixed(seed=7, regions=3)
    inp["source"]["boundingBox"] = [-100.0, -100.0, -100.0, -99.0, -99.0, -99.0]   # outside everything
    comp, src = compiled(inp, mg)
    with pytest.raises(RuntimeError, match="Infinite loop"):
        O.source_sample(comp, src, 3, 1)


@pytest.mark.parametrize("pop,b,j", [(1000, 0, 1), (1000, 3, 17), (123457, 9, 123457), (7, 1000, 3)])
def test_batch_stride_composition(pop, b, j):
    seed = 0x5EED
    s = O.rng_skip(seed, 0)                       # run: stride(getOffset) on one rank
    for _ in range(b):
        s = O.rng_skip(s, STRIDE * pop)           # after generate
        s = O.rng_skip(s, STRIDE * pop)           # after the histories
    src_stream = O.rng_skip(s, STRIDE * j)
    hist_stream = O.rng_skip(O.rng_skip(s, STRIDE * pop), STRIDE * j)
    assert src_stream == O.rng_skip(seed, STRIDE * (2 * pop * b + j))
    assert hist_stream == O.rng_skip(seed, STRIDE * (2 * pop * b + pop + j))


@pytest.mark.parametrize("top", ["transportOperatorST", "transportOperatorDT"])
def test_infinite_medium_fixed_source_closed_form(top):
    pop, cycles = 4000, 20
    inp, mg = configs.inf_medium_fixed(seed=3, nG=3, pop=pop, cycles=cycles, transport=top)
    comp, src = compiled(inp, mg)
    res, done, _ = O.fixed_run(comp, src, pop, cycles, 3, threads=4)
    assert done == cycles
    phi = configs.inf_medium_fixed_flux(mg["medium"], inp["source"]["probG"])
    expect = np.concatenate([[phi.sum()], phi]) * pop
    z = (res[:, 0] - expect) / res[:, 1]
    assert np.all(np.abs(z) < 3), z


def _absorber_slab(a, sig, top):
    mg = {"m1": {"numberOfGroups": 1, "capture": [sig], "scatteringMultiplicity": [1.0], "P0": [0.0]}}
    geom = configs._region_geometry(1, [a, 5.0, 5.0], ["m1"], [0, 0, 1, 1, 1, 1])
    tally = {"total": configs._flux_clerk()}
    src = {"type": "pointSource", "r": [0.0, 0.0, 0.0], "G": 1}
    return configs._fixed_inp(geom, mg, src, tally, 1, 1, 1, top), mg


@pytest.mark.parametrize("top", ["transportOperatorST", "transportOperatorDT"])
def test_absorbing_slab_escape_closed_form(top):
    a, sig = 1.3, 0.8
    pop, cycles = 5000, 20
    inp, mg = _absorber_slab(a, sig, top)
    comp, src = compiled(inp, mg)
    res, done, h = O.fixed_run(comp, src, pop, cycles, 11, hist=True, threads=4)
    expect = pop * (1.0 - special.expn(2, sig * a)) / sig
    z = (res[0, 0] - expect) / res[0, 1]
    assert abs(z) < 3, (res[0], expect)
    # every history either is absorbed at its one collision or leaks without one
    assert set(np.unique(h["fate"])) <= {5001, 5002}
    assert np.all((h["fate"] == 5001) == (h["n_coll"] == 1))


def test_early_stop_on_relative_sd():
    pop = 500
    inp, mg = configs.inf_medium_fixed(seed=3, nG=3, pop=pop, cycles=50)
    comp, src = compiled(inp, mg)
    full, n_full, _ = O.fixed_run(comp, src, pop, 50, 3)
    assert n_full == 50
    rel = full[0, 1] / full[0, 0]
    stop, n_stop, _ = O.fixed_run(comp, src, pop, 50, 3, stop_bin=0, stop_rel=rel * 2.0)
    assert 2 <= n_stop < 50
    assert stop[0, 1] / abs(stop[0, 0]) <= rel * 2.0


@pytest.mark.parametrize("top", ["transportOperatorST", "transportOperatorDT"])
def test_subcritical_multiplication_closed_form(top):
    """Fission secondaries in the source neutron's history (fixedSourcePhysicsPackage buffer):
    infinite medium at k_inf = 0.6, flux per source neutron (T - S - chi nuSf^T)^-1 q."""
    pop, cycles = 4000, 20
    inp, mg = configs.inf_medium_fixed_fissile(k_inf=0.6, pop=pop, cycles=cycles, transport=top)
    comp, src = compiled(inp, mg)
    k_inf, _ = configs.k_inf_and_flux(mg["medium"])
    assert abs(k_inf - 0.6) < 1e-12
    res, done, h = O.fixed_run(comp, src, pop, cycles, 1, hist=True, threads=4)
    phi = configs.fixed_flux_with_fission(mg["medium"], inp["source"]["probG"])
    expect = np.concatenate([[phi.sum()], phi]) * pop
    z = (res[:, 0] - expect) / res[:, 1]
    assert np.all(np.abs(z) < 3), z
    # without secondaries the flux would be (T - S)^-1 q: the multiplication is real
    phi0 = configs.inf_medium_fixed_flux(mg["medium"], inp["source"]["probG"])
    assert phi.sum() > 1.15 * phi0.sum()
    assert (res[0, 0] - pop * phi0.sum()) / res[0, 1] > 10
    assert h["n_coll"].max() > h["n_coll"].mean() * 5


def test_secondary_buffer_overflow_is_an_error():
    inp, mg = configs.inf_medium_fixed_fissile(k_inf=0.9, pop=200, cycles=1)
    comp, src = compiled(inp, mg)
    with pytest.raises(RuntimeError, match="Run out of space"):
        O.fixed_run(comp, src, 200, 1, 1, buffer=1)


def test_batch_problems_compile_and_run_small():
    kinds = set()
    for p in range(0, 512, 37):
        inp, mg, pop, cycles = configs.batch_problem(p, scale=0.002)
        comp, src = compiled(inp, mg)
        assert 1e3 - 1 <= comp.problem.n_bins - 1 <= 1.2e5
        assert 2 <= comp.nG <= 70 and not any(m.fissile for m in comp.mats)
        kinds.add(inp["source"]["type"])
        res, done, _ = O.fixed_run(comp, src, pop, 2, inp["seed"], threads=4)
        assert done == 2 and res[0, 0] > 0
    assert kinds == {"pointSource", "materialSource"}


# ---- neutronMGimp weight windows (neutronMGimp_class.f90 cutoffs/split/russianRoulette) ----------
def _ww(tmp_path, lower, upper, survival=2.0, **kw):
    path = tmp_path / "ww"
    path.write_text(configs.ww_file_text(lower, upper, survival))
    inp, mg = configs.ww_medium(str(path), **kw)
    return compiled(inp, mg)


def test_weight_windows_that_never_act_change_nothing(tmp_path):
    """Windows wider than any weight: no split, no roulette (whose draw is the only random number
    the windows use), so every history and score equals the plain neutronMGstd run."""
    pop, cycles = 500, 3
    comp0, src0 = compiled(*configs.ww_medium(None, pop=pop, cycles=cycles))
    comp1, src1 = _ww(tmp_path, [1e-300, 1e-300], [1e300, 1e300], pop=pop, cycles=cycles)
    assert comp1.problem.ww_n == 2 and comp1.problem.collision == T.COLL_IMP
    a, _, ha = O.fixed_run(comp0, src0, pop, cycles, 17, hist=True)
    b, _, hb = O.fixed_run(comp1, src1, pop, cycles, 17, hist=True)
    assert np.array_equal(a, b)
    assert np.array_equal(ha["rng_end"], hb["rng_end"]) and np.array_equal(ha["n_coll"], hb["n_coll"])


@pytest.mark.parametrize("top", ["transportOperatorST", "transportOperatorDT"])
def test_weight_windows_split_and_roulette_are_unbiased(tmp_path, top):
    """Active windows (split above 1.3, roulette below 0.6 to 1.2 on x < 0; split above 2.0,
    roulette below 0.8 to 1.6 on x > 0): the flux per source neutron stays the closed form
    (T - M.S)^-1 q, and the histories really split (more collisions) and roulette (some die)."""
    pop, cycles = 4000, 20
    comp0, src0 = compiled(*configs.ww_medium(None, pop=pop, cycles=cycles, transport=top))
    comp1, src1 = _ww(tmp_path, [0.6, 0.8], [1.3, 2.0], pop=pop, cycles=cycles, transport=top)
    phi = configs.ww_medium_flux()
    expect = np.concatenate([[phi.sum()], phi]) * pop
    a, _, ha = O.fixed_run(comp0, src0, pop, cycles, 17, hist=True, threads=4)
    b, _, hb = O.fixed_run(comp1, src1, pop, cycles, 17, hist=True, threads=4)
    for res in (a, b):
        z = (res[:, 0] - expect) / res[:, 1]
        assert np.all(np.abs(z) < 3.5), z
    assert hb["n_coll"].sum() > 1.1 * ha["n_coll"].sum()      # split copies are transported too
    assert not np.array_equal(ha["rng_end"], hb["rng_end"])


def test_weight_window_split_limit_and_buffer(tmp_path):
    # maxSplit caps the copies; more copies than the buffer holds is the package's error
    pop = 300
    comp, src = _ww(tmp_path, [0.01, 0.01], [0.05, 0.05], pop=pop, cycles=1, max_split=3)
    res, done, _ = O.fixed_run(comp, src, pop, 1, 17, hist=True, buffer=200)
    assert done == 1 and res[0, 0] > 0
    comp, src = _ww(tmp_path, [0.01, 0.01], [0.05, 0.05], pop=pop, cycles=1, max_split=1000)
    with pytest.raises(RuntimeError, match="Run out of space"):
        O.fixed_run(comp, src, pop, 1, 17, buffer=5)
    inp, mg = configs.ww_medium(str(tmp_path / "ww"))
    inp["type"] = "eigenPhysicsPackage"
    with pytest.raises(NotImplementedError, match="eigenvalue"):
        T.compile_problem(inp, mg)
    (tmp_path / "bad").write_text(configs.ww_file_text([1.0, 1.0], [1.5, 1.5], survival=2.0))
    inp, mg = configs.ww_medium(str(tmp_path / "bad"))
    with pytest.raises(ValueError, match="Russian Roulette survival weight"):
        T.compile_problem(inp, mg)
