"""Size-independent properties of the engine at the full sizes of the benchmark configurations
(the oracle cannot run these in seconds; the exact parity cases run at reduced size elsewhere):
  - 17x17 lattice, 7 groups, 1e6 histories per cycle, 50 + 200 cycles: delta and surface tracking
    give the same k and group-wise total flux within statistics;
  - 3D core, 70 groups, 1e7 histories per cycle, 200x200x50 mesh x 70 groups (1.4e8 bins): the mesh
    sums to a total-flux clerk to 1e-9, the k history does not depend on the tallies (bit for bit
    with a 100x100x20 mesh), and the mesh is the same with and without the extra clerk
    (shared-layout and general scoring paths) to 1e-12;
  - all 512 batched fixed-source problems: batched (10 batches per launch) against plain (one
    problem resident, one batch per launch) on every problem: identical histories and collision
    counts, results within the bound of summation order (2 N u for N scores a batch)."""
import math

import numpy as np
import pytest

from helpers import assert_results_close
from scone_amd import configs, tables as T
from scone_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu


def _eigen(inp, mg, pop, inactive, active, seed):
    comp = T.compile_problem(inp, mg)
    eng = Engine(comp, pop)
    bank = eng.fission_source(pop, seed)
    R, kh, ba, _ = eng.eigen_run(dict(pop=pop, inactive=inactive, active=active, seed=seed), bank)
    eng.close()
    return comp, R, kh, ba


def test_pin_lattice_full_size_dt_and_st_agree():
    pop, inactive, active = 1_000_000, 50, 200
    res = {}
    for top in ("transportOperatorDT", "transportOperatorST"):
        print(f"case 17x17 {top} {pop} x {inactive}+{active}", flush=True)
        inp, mg = configs.pin_lattice_7g(pop=pop, inactive=inactive, active=active, transport=top)
        res[top] = _eigen(inp, mg, pop, inactive, active, 2)
    (_, Rd, _, bd), (comp, Rs, _, bs) = res["transportOperatorDT"], res["transportOperatorST"]
    kd, ks = list(Rd.k_implicit), list(Rs.k_implicit)
    z = abs(kd[0] - ks[0]) / math.hypot(kd[1], ks[1])
    print(f"k DT {kd}, ST {ks}, z {z:.2f}", flush=True)
    assert z < 3.0
    # 7 group clerks x 289 pins x (flux, fission): group-wise total flux over the pins
    fd = bd[:, 0].reshape(7, 289, 2)[:, :, 0].sum(axis=1)
    fs = bs[:, 0].reshape(7, 289, 2)[:, :, 0].sum(axis=1)
    sd = np.sqrt((bd[:, 1].reshape(7, 289, 2)[:, :, 0] ** 2).sum(axis=1) +
                 (bs[:, 1].reshape(7, 289, 2)[:, :, 0] ** 2).sum(axis=1))
    print(f"group flux DT {fd}, ST {fs}", flush=True)
    assert np.all(np.abs(fd - fs) <= np.maximum(5.0 * sd, 2e-3 * np.abs(fd)))


def test_core_3d_full_mesh_sums_to_total_and_k_independent_of_tallies():
    pop, inactive, active, seed = 10_000_000, 3, 5, 5
    print("case core 3D, 1.4e8-bin mesh", flush=True)
    inp, mg = configs.core_3d(pop=pop, inactive=inactive, active=active)
    comp, R1, kh1, b1 = _eigen(inp, mg, pop, inactive, active, seed)
    assert comp.problem.n_bins == 200 * 200 * 50 * 70
    print("case core 3D, mesh + total clerk (general scoring path)", flush=True)
    inp2, mg2 = configs.core_3d(pop=pop, inactive=inactive, active=active)
    inp2["activeTally"]["total"] = {"type": "collisionClerk", "response": ["flux"], "flux": {"type": "fluxResponse"}}
    comp2, R2, kh2, b2 = _eigen(inp2, mg2, pop, inactive, active, seed)
    assert np.array_equal(kh1.view(np.uint64), kh2.view(np.uint64))
    n_mesh = 200 * 200 * 50 * 70
    off = {c["name"]: c["addr"] for c in comp2.clerk_info}
    total = b2[off["total"], 0]
    mesh2 = np.concatenate([b2[off[f"mesh_g{g}"]:off[f"mesh_g{g}"] + 200 * 200 * 50, 0] for g in range(1, 71)])
    assert mesh2.size == n_mesh
    rel = abs(mesh2.sum() - total) / total
    print(f"mesh sum {mesh2.sum():.12e}, total clerk {total:.12e}, rel {rel:.2e}", flush=True)
    assert rel < 1e-9
    mesh1 = b1[:, 0]
    scale = np.maximum(np.abs(mesh1), np.abs(mesh2))
    assert np.all(np.abs(mesh1 - mesh2) <= 1e-12 * scale)
    print("case core 3D, 100x100x20 mesh", flush=True)
    inp3, mg3 = configs.core_3d(pop=pop, inactive=inactive, active=active, mesh=(100, 100, 20))
    _, R3, kh3, b3 = _eigen(inp3, mg3, pop, inactive, active, seed)
    assert np.array_equal(kh1.view(np.uint64), kh3.view(np.uint64))
    assert list(R1.k_implicit) == list(R3.k_implicit)
    # the coarse mesh holds the same total flux (every collision lands in both meshes)
    assert abs(b3[:, 0].sum() - mesh1.sum()) <= 1e-9 * mesh1.sum()


def test_all_512_batched_problems_equal_plain():
    n = 512
    print(f"case batched vs plain, {n} problems (compiling)", flush=True)
    probs = []
    for p in range(n):
        inp, mg, pop, cycles = configs.batch_problem(p, transport="transportOperatorDT")
        comp = T.compile_problem(inp, mg)
        probs.append((comp, T.compile_source(inp["source"], comp), pop, cycles, inp["seed"]))
    budget = 0
    for (comp, _, pop, _, _) in probs:
        budget += 2 * pop * 200 + 8 * 12 * int(comp.problem.n_bins) + (1 << 20)
    budget = int(budget * 1.25) + (256 << 20)
    print("case batched run", flush=True)
    # the launch shape the batched benchmark's autotuner picks (tile 1024, refill 16)
    B = Batch(transport=T.TRANSPORT_DT, max_slots=n, hbm_budget=budget, use_graph=1, tile=1024, n_launch=1, refill=16)
    slots = [B.add(c, s, pop, cyc, seed, window=10) for (c, s, pop, cyc, seed) in probs]
    B.run_all()
    batched = [B.query(s) for s in slots]
    B.close()
    print("case plain runs", flush=True)
    P = Batch(transport=T.TRANSPORT_DT, max_slots=1, hbm_budget=budget, use_graph=0, tile=256, n_launch=1)
    for i, (c, s, pop, cyc, seed) in enumerate(probs):
        slot = P.add(c, s, pop, cyc, seed, window=1)
        P.run_all()
        r, bins, _, _ = P.query(slot)
        rb, bb, _, _ = batched[i]
        assert r.n_batches == rb.n_batches == cyc
        assert r.n_histories == rb.n_histories and r.n_collisions == rb.n_collisions, i
        # the same scores summed by FP64 atomics in two orders: each sum of N positive terms is
        # within (N - 1) u of the exact one (u = 2^-53), N at most the collisions of a batch
        n_per_batch = r.n_collisions / cyc
        rtol = max(1e-12, 2.0 * n_per_batch * 2.0 ** -53)
        assert_results_close(bins, bb, cyc, rtol=rtol, label=f"problem {i}")
        P.release(slot)
        if i % 128 == 127:
            print(f"  {i + 1} problems compared", flush=True)
    P.close()
