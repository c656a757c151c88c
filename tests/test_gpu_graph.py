"""Captured-graph split steps (bec_split_step_graph) and the tile autotuner.

A graph replays the launches that bec_split_step issues, so its results must be bitwise
those of the eager path, including after parameter changes between replays (the kernels
read parameters from device memory) and after the active set changes (a new capture).
Each replay is checked against the eager workspace with eager device work (energies,
copies) in between."""
import numpy as np
import pytest

import becgpu as G
from becgpu import api as A
from becgpu.serve import Server
from helpers import random_psi
from test_gpu_serve import _mix

pytestmark = pytest.mark.gpu


def _ws(case, precision):
    if case == "itp2d":
        grid = G.make_grid(G.GridConfig((32, 32), (20.0, 20.0)))
        F, itp, c1, ddi = 1, True, -0.5, False
    elif case == "rt1d_spin6":
        grid = G.make_grid(G.GridConfig((64,), (20.0,)))
        F, itp, c1, ddi = 6, False, 0.4, False
    else:  # rt2d_spin6_ddi
        grid = G.make_grid(G.GridConfig((32, 32), (12.0, 12.0)))
        F, itp, c1, ddi = 6, False, 0.3, True
    D = 2 * F + 1
    sp = G.SimParams(dt=0.002, n_steps=10, imaginary_time=itp)
    psi = random_psi(grid.shape, D, 3, scale=0.05)
    ws = A.make_workspace(grid=grid, atom=G.AtomSpecies("t", 1.0, F),
                          interactions=G.InteractionParams(8.0, c1), potential=G.HarmonicTrap(*([1.0] * grid.ndim)),
                          sim_params=sp, psi_init=psi, enable_ddi=ddi, c_dd=0.5 if ddi else float("nan"),
                          precision=precision)
    return ws


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", ["itp2d", "rt1d_spin6", "rt2d_spin6_ddi"])
def test_graph_replays_equal_eager(case, precision):
    a = _ws(case, precision)
    b = _ws(case, precision)
    for rep in range(5):
        if rep == 2:  # new couplings between replays: read from device memory by the graph
            for w in (a, b):
                w.set_instance(0, interactions=G.InteractionParams(11.0, 0.2),
                               potential=G.HarmonicTrap(*([1.0] * w.grid.ndim)),
                               c_dd=0.5 if w.enable_ddi else 0.0)
        A.split_step(a, 7)
        A.split_step(b, 7, graph=True)
        ea, eb = a.energy_terms(), b.energy_terms()       # eager work between replays
        assert np.array_equal(ea, eb), (rep, ea, eb)
        assert np.array_equal(a.get_psi(0), b.get_psi(0)), rep
    a.close()
    b.close()


def test_graph_active_set_and_operator_changes_recapture():
    grid = G.make_grid(G.GridConfig((32, 32), (20.0, 20.0)))
    sp = G.SimParams(dt=0.002, n_steps=10, imaginary_time=True)
    ws = []
    for _ in range(2):
        w = G.Workspace(grid=grid, F=1, n_instances=2, sim_params=sp)
        for i in range(2):
            w.set_instance(i, interactions=G.InteractionParams(8.0, -0.5 + i),
                           potential=G.HarmonicTrap(1.0, 1.0))
            w.set_psi(random_psi(grid.shape, 3, 20 + i, scale=0.05), i)
        ws.append(w)
    a, b = ws

    def run(n):
        A.split_step(a, n)
        A.split_step(b, n, graph=True)
        for i in a.active:
            assert np.array_equal(a.get_psi(i), b.get_psi(i))

    run(4)
    run(4)
    for w in ws:
        w.set_active([1])
    run(4)
    for w in ws:  # spin mixing off on the active instance: another schedule
        w.set_instance(1, interactions=G.InteractionParams(8.0, 0.0), potential=G.HarmonicTrap(1.0, 1.0))
    run(4)
    for w in ws:
        w.set_active([0, 1])
    run(4)
    run(4)
    for w in ws:
        w.close()


def test_server_with_graphs_equals_eager():
    reqs = _mix("fp32")[:8]
    r0, _ = Server(slots_per_pool=2, chunk_ms=0.5, max_chunk=32, keep_psi=True).run(reqs, realtime=False)
    r1, st = Server(slots_per_pool=2, chunk_ms=0.5, max_chunk=32, keep_psi=True,
                    graphs=True).run(reqs, realtime=False)
    print(st)
    by0 = {r.rid: r for r in r0}
    for r in r1:
        q = by0[r.rid]
        assert (r.steps, r.converged, r.energy) == (q.steps, q.converged, q.energy)
        assert np.array_equal(r.psi, q.psi)


def test_autotune_is_per_shape_and_leaves_psi():
    grid = G.make_grid(G.GridConfig((64, 64, 64), (20.0,) * 3))
    sp = G.SimParams(dt=0.001, n_steps=1, imaginary_time=False)
    w1 = G.Workspace(grid=grid, F=1, n_instances=1, sim_params=sp, autotune=False)
    psi = random_psi(grid.shape, 3, 9)
    w1.set_psi(psi, 0)
    r1 = w1.autotune()
    assert r1 >= 1 and 64 * 64 % r1 == 0 and 3 * r1 * 64 // 8 <= 512
    assert np.array_equal(w1.get_psi(0), psi)
    w2 = G.Workspace(grid=grid, F=1, n_instances=3, sim_params=sp)
    assert w2.xrows == r1
    w1.close()
    w2.close()
