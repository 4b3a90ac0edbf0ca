"""MG data, tally arithmetic and population control of the oracle, pinned (CPU).

* baseMgNeutronDatabase_iTest.f90:88-131 known answers on IntegrationTestFiles/mgMat1, mgMat2
  (copied to tests/golden as data).
* scoreMemory_test.f90:92-150 known answers (reference LCG scores, closeCycle with norm 0.7).
* keffImplicitClerk_test.f90 known answer k = 0.906521739130435 +/- 0.006521739130435.
* particleDungeon_test.f90:209-374 sizes and brood ordering for normSize_Repr, plus the invariants
  that define it: exactly totPop survivors, the dropped / duplicated sites are those whose
  skip-ahead random numbers are the smallest, brood order is stable.
"""
import math
import os

import numpy as np
import pytest

import binding as O
from helpers import geometry_problem
from scone_amd import dictio, tables as T

HERE = os.path.dirname(os.path.abspath(__file__))


def mg_db(active=("mat1",)):
    mg = {"mat1": dictio.read(os.path.join(HERE, "golden", "mgMat1")),
          "mat2": dictio.read(os.path.join(HERE, "golden", "mgMat2"))}
    unis = {"root": {"id": 1, "type": "rootUniverse", "border": 1, "fill": active[0]}}
    surfs = {"b": {"id": 1, "type": "box", "origin": [0.0, 0.0, 0.0], "halfwidth": [1.0, 1.0, 1.0]}}
    inp = {"transportOperator": {"type": "transportOperatorST"},
           "geometry": {"type": "geometryStd", "boundary": [1] * 6, "graph": {"type": "shrunk"},
                        "surfaces": surfs, "cells": {}, "universes": unis},
           "nuclearData": {"handles": {"mg": {"type": "baseMgNeutronDatabase", "PN": "P0"}},
                           "materials": {"mat1": {"xsFile": "x"}, "mat2": {"xsFile": "y"}}},
           "activeTally": {}}
    return T.compile_problem(inp, mg)


def test_mg_database_known_answers():
    comp = mg_db()
    TOL = 1e-6
    xs = comp.xs   # [mat][G][TOTAL, SCATTER, CAPTURE, FISSION, NUFISSION, KAPPA]
    assert comp.nG == 4
    assert xs[0, 0, 0] == pytest.approx(2.1, abs=TOL)          # getTrackingXS(G=1, mat1)
    assert xs[1, 0, 0] == pytest.approx(3.1, abs=TOL)          # getTotalMatXS(G=1, mat2)
    assert xs[0, 2, 0] == pytest.approx(6.0, abs=TOL)          # G=3, mat1
    assert comp.majorant[0] == pytest.approx(2.1, abs=TOL)     # activate([1])
    assert list(xs[1, 0, :]) == pytest.approx([3.1, 1.1, 1.0, 1.0, 2.3, 202.0], abs=TOL)
    assert list(xs[0, 3, :]) == pytest.approx([7.1, 3.1, 4.0, 0.0, 0.0, 0.0], abs=TOL)
    assert comp.problem.mg.fissile[0] == 0 and comp.problem.mg.fissile[1] == 1


def test_mg_sampling_cdf_brute_force():
    """multiScatterMG%sampleGout and fissionMG chi scan equal the inverse CDF on a fine u grid."""
    comp = mg_db()
    P0 = np.array(comp.mats[1].P0).reshape(4, 4)    # [gin][gout]
    for gin in range(4):
        s = comp.xs[1, gin, 1]
        cdf = np.cumsum(P0[gin]) / s
        for u in np.linspace(0.0, 0.999999, 997):
            rem = u * s
            g = None
            for gout in range(4):
                rem = rem - P0[gin, gout]
                if rem < 0.0:
                    g = gout
                    break
            assert g == int(np.searchsorted(cdf, u, side="right")) or abs(cdf[g] - u) < 1e-12


def test_score_memory_known_answers():
    # scoreMemory_test.f90: reference LCG (A 2469, M 65521, seed 9294) scores, 10 cycles of 20
    A, M = 2469, 65521
    seed = 9294
    rnd = []
    for _ in range(200):
        seed = (A * seed) % M
        rnd.append(seed / float(M))
    pi = 4.0 * math.atan(1.0)
    scores = [2.0 + math.sin(pi * r - pi / 2) for r in rnd]
    scoresInt = [int(r * 100) for r in rnd]
    bins = np.zeros(2)
    csum = np.zeros(2)
    csum2 = np.zeros(2)
    for i in range(10):
        for j in range(20 * i, 20 * (i + 1)):
            bins[0] = bins[0] + scores[j]
            bins[1] = bins[1] + float(scoresInt[j])
        O.close_cycle(bins, csum, csum2, 0.7)
    m, s = O.score_result(csum[0], csum2[0], 10)
    assert m == pytest.approx(26.401471259728442, abs=1e-9)
    assert s == pytest.approx(0.645969443981583, abs=1e-9)
    m, s = O.score_result(csum[1], csum2[1], 10)
    assert m == pytest.approx(623.0, abs=1e-9)
    assert s == pytest.approx(27.982494527829360, abs=1e-9)
    # accumulate-only bins with 200 samples (bins 4 and 5)
    c1 = sum(scores)
    c2 = sum(x * x for x in scores)
    m, s = O.score_result(c1, c2, 200)
    assert m == pytest.approx(1.885819375694888, abs=1e-9)
    assert s == pytest.approx(0.049102082638055, abs=1e-9)
    c1 = float(sum(scoresInt))
    c2 = float(sum(x * x for x in scoresInt))
    m, s = O.score_result(c1, c2, 200)
    assert m == pytest.approx(44.5, abs=1e-9)
    assert s == pytest.approx(2.015580019267494, abs=1e-9)


def test_keff_implicit_known_answer():
    # keffImplicitClerk_test.f90: mock XS total 1, capture 2, fission 1, nuFission 3;
    # (IMP_PROD, SCATTER_PROD, IMP_ABS, ANA_LEAK) per cycle
    k1 = O.keff_implicit([3.0 * 0.7, 0.1, 3.0 * 0.7, 0.3])
    k2 = O.keff_implicit([3.0 * 0.6, 0.1, 3.0 * 0.6, 0.3])
    m, s = O.score_result(k1 + k2, k1 * k1 + k2 * k2, 2)
    assert m == pytest.approx(0.906521739130435, abs=1e-9)
    assert s == pytest.approx(0.006521739130435, abs=1e-9)


def bank_with_broods(broods, seed=0):
    b = T.bank_array(len(broods))
    b["broodID"] = broods
    b["wgt"] = 1.0
    b["G"] = np.arange(len(broods)) + 1
    b["r"][:, 0] = np.arange(len(broods))
    return b


@pytest.mark.parametrize("tot", [5, 8, 10, 15, 23, 31])
def test_norm_size_repr_invariants(tot):
    broods = [3, 1, 2, 1, 3, 2, 4, 4, 1, 2]
    b = bank_with_broods(broods)
    master = 0x1234567
    out = O.norm_size_repr(b, tot, master)
    assert len(out) == tot
    assert list(out["broodID"]) == sorted(out["broodID"])       # brood-sorted
    # the reference's draw for the j-th site (sorted order) is the j-th number after the master
    srt = b[np.argsort(b["broodID"], kind="stable")]
    n = len(srt)
    rn = []
    s = master
    for _ in range(n):
        u, s = O.rng_get(s)
        rn.append(u)
    rn = np.array(rn)
    excess = n - tot
    if excess > 0:
        thr = np.sort(rn)[excess - 1]
        kept = srt[rn > thr]
        assert np.array_equal(out["G"], kept["G"])
    elif excess < 0:
        n_copies = (-excess) // n
        n_dup = (-excess) % n
        counts = {g: n_copies + 1 for g in srt["G"]}
        if n_dup:
            thr = np.sort(rn)[n_dup - 1]
            for g in srt["G"][rn <= thr]:
                counts[g] += 1
        got = {}
        for g in out["G"]:
            got[g] = got.get(g, 0) + 1
        assert got == counts


def test_sort_by_brood_with_duplicates():
    # particleDungeon_test.f90:340-374: sites of each brood stay in emission order
    broods = [2, 1, 2, 3, 1, 3, 2]
    b = bank_with_broods(broods)
    out = O.norm_size_repr(b, len(broods), 5)       # excess 0: sort only
    assert list(out["broodID"]) == [1, 1, 2, 2, 2, 3, 3]
    assert list(out["G"]) == [2, 5, 1, 3, 7, 4, 6]


def test_neutron_mg_imp_sites_keep_the_collision_count():
    """neutronMGimp without weight windows differs from neutronMGstd only in the fission sites'
    collisionN (neutronMGimp_class.f90 implicit keeps the colliding particle's count): same RNG
    use, same bank otherwise, same scores."""
    import binding as O
    from scone_amd import configs, tables as T
    pop = 800
    inp, mg = configs.pin_lattice_7g(pop=pop, transport="transportOperatorDT")
    std = T.compile_problem(inp, mg)
    inp["collisionOperator"] = {"neutronMG": {"type": "neutronMGimp", "maxSplit": 50}}
    imp = T.compile_problem(inp, mg)
    assert imp.problem.collision == T.COLL_IMP and imp.problem.max_split == 50
    bank = O.fission_source(std, pop, 77)
    a, ab, aks, _ = O.eigen_cycle(std, bank, 99, 1.0)
    b, bb, bks, _ = O.eigen_cycle(imp, bank, 99, 1.0)
    assert len(a) == len(b)
    for f in a.dtype.names:
        if f != "collisionN":
            assert np.array_equal(a[f], b[f]), f
    assert np.all(a["collisionN"] == 0)
    assert b["collisionN"].min() >= 0 and b["collisionN"].max() > 3
    assert np.array_equal(ab, bb) and np.array_equal(aks, bks)
    # sites of one history come in emission order (stable broodID sort), so their counts do not
    # decrease; the history's first collision gives 0
    for br in np.unique(b["broodID"]):
        cn = b["collisionN"][b["broodID"] == br]
        assert np.all(np.diff(cn) >= 0), br
    assert (b["collisionN"] == 0).any()
    inp["collisionOperator"]["neutronMG"]["weightWindows"] = 1
    with pytest.raises(NotImplementedError):
        T.compile_problem(inp, mg)
