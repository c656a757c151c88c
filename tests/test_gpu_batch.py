"""Batched ground states (becgpu.batch): every solve of a mixed-shape batch must equal the
same solve run alone through find_ground_state (identical arithmetic per instance), and the
FP64 batch must match the oracle."""
import math

import numpy as np
import pytest

from oracle import spinor_oracle as O
import becgpu as G
from becgpu import api as A
from becgpu.batch import ITPJob, multistart_batched, solve_batch, solve_batch_refined
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def _jobs():
    grids = {n: G.make_grid(G.GridConfig((n, n), (20.0, 20.0))) for n in (32, 64)}
    kinds = ("polar", "ferromagnetic", "uniform", "antiferromagnetic", "random")
    jobs = []
    for k in range(7):
        n = 32 if k % 3 else 64
        psi0 = G.init_psi(grids[n], G.SpinSystem(1), kinds[k % 5], seed=100 + k)
        jobs.append(ITPJob(grid=grids[n], F=1, interactions=G.InteractionParams(10.0, -1 + 0.3 * k),
                           psi0=psi0, potential=G.HarmonicTrap(1.0, 1.0), dt=0.005, n_steps=400,
                           tol=1e-7))
    return jobs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_equals_single_solves(precision):
    jobs = _jobs()
    res, stats = solve_batch(jobs, precision=precision, keep_psi=True)
    print(stats)
    assert stats["n_groups"] == 2
    for j, r in zip(jobs, res):
        s = A.find_ground_state(grid=j.grid, atom=G.Rb87, interactions=j.interactions,
                                potential=j.potential, dt=j.dt, n_steps=j.n_steps, tol=j.tol,
                                psi_init=j.psi0, precision=precision)
        assert r["steps"] == s["steps"] and r["converged"] == s["converged"]
        assert r["energy"] == s["energy"]
        assert np.array_equal(r["psi"], s["workspace"].get_psi())


def test_batch_fp64_matches_oracle():
    jobs = _jobs()[:3]
    res, _ = solve_batch(jobs, precision="fp64", keep_psi=True)
    for j, r in zip(jobs, res):
        og = O.make_grid(j.grid.config.n_points, j.grid.config.box_size)
        ows = O.make_workspace(og, 1, j.psi0, j.dt, True, omega=(1.0, 1.0), c0=10.0,
                               c1=j.interactions.c1, n_steps=j.n_steps)
        ro = O.find_ground_state(ows, j.n_steps, j.tol)
        assert r["steps"] == ro["steps"]
        assert abs(r["energy"] - ro["energy"]) <= 1e-12 * abs(ro["energy"])
        assert rel_l2(r["psi"], ows.psi) <= 1e-10


def test_multistart_batched_matches_sequential():
    grid = G.make_grid(G.GridConfig((32, 32), (20.0, 20.0)))
    kw = dict(grid=grid, atom=G.Rb87, interactions=G.InteractionParams(10.0, -0.5),
              potential=G.HarmonicTrap(1.0, 1.0), dt=0.005, n_steps=300, tol=1e-7)
    b = multistart_batched(initial_states=("polar", "ferromagnetic", "random"), n_random=2, **kw)
    s = A.find_ground_state_multistart(initial_states=("polar", "ferromagnetic", "random"),
                                       n_random=2, **kw)
    assert b["initial_state"] == s["initial_state"]
    assert b["energy"] == s["energy"]
    assert [r["energy"] for r in b["all_results"]] == [r["energy"] for r in s["all_results"]]


def test_continuation_chains_batched_equal_sequential():
    """Three continuation chains (c1 sweeps at different c0) through the batched path equal
    scan_continuation chain by chain, bit for bit, fallbacks included."""
    from becgpu.batch import scan_continuation_batched
    grid = G.make_grid(G.GridConfig((16,), (10.0,)))
    chains = [([-1.0, 0.5, 1.0], lambda v: G.InteractionParams(10.0, v)),
              ([0.2, 0.3], lambda v: G.InteractionParams(6.0, v)),
              ([10.0, 10.5, 20.0], lambda v: G.InteractionParams(v, -0.5))]
    kw = dict(grid=grid, atom=G.Rb87, initial_state="polar", n_steps_continuation=300,
              n_steps_fresh=600, dt=0.01, tol=1e-8, potential=G.HarmonicTrap(1.0))
    got = scan_continuation_batched(chains, **kw)
    for (vals, mk), rec in zip(chains, got):
        ref = A.scan_continuation(param_values=vals, make_interactions=mk, **kw)
        assert len(rec) == len(ref) == len(vals)
        for g, r in zip(rec, ref):
            assert (g["param"], g["converged"], g["phase"], g["fallback"]) == \
                (r["param"], r["converged"], r["phase"], r["fallback"])
            assert g["energy"] == r["energy"]
            assert np.array_equal(g["psi"], r["psi"])
    assert any(r["fallback"] for r in got[2])


def test_refined_fp32_batch_reaches_fp64_fixed_point():
    """batch.solve_batch_refined: an untrapped 2-D box (uniform ground state) relaxed in
    FP32-bulk keeps its slowest mode frozen (the per-step decay rounds away); the FP64 steps
    that finish the solve must bring the energy to the FP64 fixed point's within 1e-7
    (relative), and each refined result must equal an FP64 batch started from the FP32
    state."""
    grid = G.make_grid(G.GridConfig((32, 32), (20.0, 20.0)))
    jobs = []
    for k, kind in enumerate(("polar", "ferromagnetic")):
        psi0 = G.init_psi(grid, G.SpinSystem(1), kind, seed=7 + k)
        jobs.append(ITPJob(grid=grid, F=1, interactions=G.InteractionParams(10.0, -0.5),
                           psi0=psi0, dt=0.005, n_steps=40000, tol=0.0))
    r32, _ = solve_batch(jobs, precision="fp32", keep_psi=True)
    rr, stats = solve_batch_refined(jobs, refine_steps=3000, keep_psi=True)
    r64, _ = solve_batch(jobs, precision="fp64")
    for a, b, c in zip(r32, rr, r64):
        print("fp32", abs(a["energy"] - c["energy"]) / abs(c["energy"]),
              "refined", abs(b["energy"] - c["energy"]) / abs(c["energy"]))
        assert abs(b["energy"] - c["energy"]) <= 1e-7 * abs(c["energy"])
        assert b["steps_fp32"] == a["steps"] and b["energy_fp32"] == a["energy"]
    again = [ITPJob(grid=j.grid, F=1, interactions=j.interactions, psi0=a["psi"], dt=j.dt,
                    n_steps=3000, tol=0.0) for j, a in zip(jobs, r32)]
    r2, _ = solve_batch(again, precision="fp64", keep_psi=True)
    for b, c in zip(rr, r2):
        assert b["energy"] == c["energy"]
        assert np.array_equal(b["psi"], c["psi"])
