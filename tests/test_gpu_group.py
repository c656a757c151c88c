"""Grouped split steps (bec_group_*): workspaces of different grid shapes advanced with one
kernel launch per pass over all of them, against each workspace advanced on its own.

The reference solves its batches one at a time (find_ground_state_multistart and
scan_continuation loop over find_ground_state, ground_state.jl:239-271, 328-378); the
grouped launch must not change any result, so every case compares psi bit for bit with
the same workspaces advanced by bec_split_step, and the batch driver's grouped mode with
its per-shape mode.
"""
import numpy as np
import pytest

import becgpu as G
from becgpu.batch import ITPJob, WorkspaceGroup, solve_batch
from becgpu.physics import cell_volume
from helpers import random_psi

pytestmark = pytest.mark.gpu


def _ws(n, F, itp, precision, seed, n_inst=2, mixing=True, autotune=False):
    grid = G.make_grid(G.GridConfig(tuple(n), tuple(12.0 for _ in n)))
    ws = G.Workspace(grid=grid, F=F, n_instances=n_inst, precision=precision,
                     sim_params=G.SimParams(dt=0.004, n_steps=100, imaginary_time=itp),
                     autotune=autotune)
    for i in range(n_inst):
        c1 = -0.4 + 0.3 * i if mixing else 0.0
        ws.set_instance(i, interactions=G.InteractionParams(8.0 + 2 * i, c1),
                        zeeman=G.ZeemanParams(0.1 * i, 0.05),
                        potential=G.HarmonicTrap(*(1.0 + 0.1 * d for d in range(len(n)))))
        psi = random_psi(grid.shape, 2 * F + 1, seed + i)
        ws.set_psi(psi / np.sqrt(np.sum(np.abs(psi) ** 2) * cell_volume(grid)), i)
    return ws


def _pair(shapes, F, itp, precision, mixing=True, autotune=False):
    """Two identical sets of workspaces; default x tiles (one block size for every shape,
    as grouped x-pass launches need) unless autotune."""
    a = [_ws(n, F, itp, precision, 10 * k, mixing=mixing, autotune=autotune)
         for k, n in enumerate(shapes)]
    b = [_ws(n, F, itp, precision, 10 * k, mixing=mixing, autotune=autotune)
         for k, n in enumerate(shapes)]
    return a, b


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("F,itp", [(1, True), (2, False), (0, True)])
def test_grouped_2d_equals_separate(precision, graph, F, itp):
    """Spin 1 with spin mixing (ITP), spin 2 without it (real time: its spin mixing runs in
    the lane-per-site pass, which has no grouped kernel), the scalar gas."""
    shapes = [(32, 32), (64, 64), (128, 128), (256, 32)]
    a, b = _pair(shapes, F, itp, precision, mixing=F != 2)
    grp = WorkspaceGroup(a)
    try:
        for k in (1, 3, 7, 3):
            grp.split_step(k, graph=graph)
            assert grp.grouped
            for w in b:
                G.api.split_step(w, k)
        for wa, wb in zip(a, b):
            for i in range(2):
                assert np.array_equal(wa.get_psi(i), wb.get_psi(i))
        # an instance retired from one member: the next launches follow the new active set
        a[1].set_active([1])
        b[1].set_active([1])
        grp.split_step(4, graph=graph)
        for w in b:
            G.api.split_step(w, 4)
        for wa, wb in zip(a, b):
            for i in range(2):
                assert np.array_equal(wa.get_psi(i), wb.get_psi(i))
    finally:
        grp.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_grouped_3d_equals_separate(precision):
    shapes = [(32, 32, 32), (64, 64, 32), (128, 32, 64)]
    a, b = _pair(shapes, 1, True, precision)
    grp = WorkspaceGroup(a)
    try:
        grp.split_step(5, graph=True)
        assert grp.grouped
        for w in b:
            G.api.split_step(w, 5)
        for wa, wb in zip(a, b):
            assert np.array_equal(wa.get_psi(0), wb.get_psi(0))
            assert np.array_equal(wa.get_psi(1), wb.get_psi(1))
    finally:
        grp.close()


def test_shapes_without_grouped_kernels_fall_back():
    """x side 48 (and 96-point lines) have no grouped kernel, nor has spin 2 with spin
    mixing, nor have members with different x-pass block sizes (autotuned tiles): the
    members are advanced one at a time, with the same results."""
    shapes = [(48, 48), (64, 96)]
    a, b = _pair(shapes, 1, True, "fp64")
    grp = WorkspaceGroup(a)
    try:
        grp.split_step(6, graph=True)
        assert not grp.grouped
        for w in b:
            G.api.split_step(w, 6)
        for wa, wb in zip(a, b):
            assert np.array_equal(wa.get_psi(0), wb.get_psi(0))
    finally:
        grp.close()
    a, b = _pair([(32, 32), (64, 64)], 2, False, "fp64")
    grp = WorkspaceGroup(a)
    try:
        grp.split_step(3, graph=False)
        assert not grp.grouped
        for w in b:
            G.api.split_step(w, 3)
        for wa, wb in zip(a, b):
            assert np.array_equal(wa.get_psi(1), wb.get_psi(1))
    finally:
        grp.close()
    a, b = _pair([(32, 32), (256, 256)], 1, True, "fp32", autotune=True)
    grp = WorkspaceGroup(a)
    try:
        grp.split_step(3, graph=True)
        if a[0].xrows * 32 != a[1].xrows * 256:
            assert not grp.grouped
        for w in b:
            G.api.split_step(w, 3)
        for wa, wb in zip(a, b):
            assert np.array_equal(wa.get_psi(1), wb.get_psi(1))
    finally:
        grp.close()


def test_group_rejects_mixed_members():
    g1 = G.make_grid(G.GridConfig((32, 32), (10.0, 10.0)))
    a = G.Workspace(grid=g1, F=1, n_instances=1, precision="fp64",
                    sim_params=G.SimParams(dt=0.01, n_steps=10, imaginary_time=True))
    b = G.Workspace(grid=g1, F=2, n_instances=1, precision="fp64",
                    sim_params=G.SimParams(dt=0.01, n_steps=10, imaginary_time=True))
    with pytest.raises(Exception):
        WorkspaceGroup([a, b])
    with pytest.raises(Exception):
        WorkspaceGroup([a, a])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_solve_batch_grouped_equals_per_shape(precision):
    """Ground-state batch over 32^2, 64^2 and 128^2 (cfg4's pattern at test size): grouped
    launches (captured), per-shape graphs and plain per-shape launches with the same tiles
    agree bit for bit."""
    jobs = []
    states = ("polar", "ferromagnetic", "uniform", "antiferromagnetic")
    for k in range(9):
        n = (32, 64, 128)[k % 3]
        grid = G.make_grid(G.GridConfig((n, n), (20.0, 20.0)))
        psi = G.init_psi(grid, G.SpinSystem(1), states[k % 4])
        jobs.append(ITPJob(grid=grid, F=1, interactions=G.InteractionParams(10.0, -1.0 + 0.25 * k),
                           psi0=psi, potential=G.HarmonicTrap(1.0, 1.0), dt=0.005, n_steps=300,
                           tol=1e-8))
    ra, sa = solve_batch(jobs, precision=precision, keep_psi=True, graphs=True, grouped=True)
    rb, _ = solve_batch(jobs, precision=precision, keep_psi=True, graphs=True, autotune=False)
    rc, _ = solve_batch(jobs, precision=precision, keep_psi=True, autotune=False)
    assert sa["grouped_intervals"] > 0
    for x, y, z in zip(ra, rb, rc):
        assert x["steps"] == y["steps"] == z["steps"]
        assert x["converged"] == y["converged"] == z["converged"]
        assert x["energy"] == y["energy"] == z["energy"]
        assert np.array_equal(x["psi"], y["psi"]) and np.array_equal(x["psi"], z["psi"])
