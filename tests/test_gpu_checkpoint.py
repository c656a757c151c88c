"""run_simulation_checkpointed! on the device (test_checkpoint.jl:1-80): files are written,
a resumed run equals the continuous one, the callback cadence is kept."""
import os

import numpy as np
import pytest

import becgpu as G
from becgpu import api as A
from becgpu.io import load_state, run_simulation_checkpointed

pytestmark = pytest.mark.gpu

GRID = G.make_grid(G.GridConfig((16,), (10.0,)))
DT = 0.01


def _ws(n_steps, save_every, psi=None):
    sp = G.SimParams(dt=DT, n_steps=n_steps, save_every=save_every)
    return A.make_workspace(grid=GRID, atom=G.Rb87, interactions=G.InteractionParams(1.0, 0.0),
                            sim_params=sp, psi_init=psi)


def test_checkpoint_creates_files(tmp_path):
    ws = _ws(100, 50)
    run_simulation_checkpointed(ws, checkpoint_dir=str(tmp_path), checkpoint_every=50)
    files = sorted(os.listdir(tmp_path))
    assert files == ["step_00000050.npz", "step_00000100.npz"]
    s = load_state(os.path.join(tmp_path, files[-1]))
    assert s["step"] == 100 and s["t"] == pytest.approx(1.0)
    assert np.array_equal(s["psi"], ws.get_psi(0))
    assert (s["c0"], s["c1"], s["dt"], s["imaginary_time"]) == (1.0, 0.0, DT, False)


def test_resume_equals_continuous(tmp_path):
    full = _ws(100, 100)
    psi0 = full.get_psi(0)
    A.run_simulation(full)
    first = _ws(50, 50, psi0)
    run_simulation_checkpointed(first, checkpoint_dir=str(tmp_path), checkpoint_every=50)
    second = _ws(100, 50, psi0)                  # resume=True restores psi, t, step
    run_simulation_checkpointed(second, checkpoint_dir=str(tmp_path), checkpoint_every=50,
                                resume=True)
    assert int(second.step[0]) == 100 and second.t[0] == pytest.approx(1.0)
    assert np.max(np.abs(second.get_psi(0) - full.get_psi(0))) < 1e-10


def test_checkpoint_callback_cadence(tmp_path):
    ws = _ws(50, 25)
    count = []
    run_simulation_checkpointed(ws, checkpoint_dir=str(tmp_path), checkpoint_every=25,
                                callback=lambda w, s: count.append(s))
    assert count == [25, 50]


def test_resumed_run_matches_oracle(tmp_path):
    """A real-time run interrupted at step 40 and resumed from its checkpoint
    (run_simulation_checkpointed!, simulation.jl:96-148) equals the oracle's uninterrupted
    run_simulation! with the same save points (leapfrog, simulation.jl:49-94): psi to 1e-10
    and the energy of the final state to 1e-12."""
    from oracle import spinor_oracle as O
    n, L = (32,), (10.0,)
    og = O.make_grid(n, L)
    grid = G.make_grid(G.GridConfig(n, L))
    psi0 = G.init_psi(grid, G.SpinSystem(1), "uniform")
    psi0[0] *= np.exp(0.4j * grid.x[0])                       # a spin texture that moves

    def ws(n_steps):
        sp = G.SimParams(dt=DT, n_steps=n_steps, save_every=20)
        return A.make_workspace(grid=grid, atom=G.Rb87, interactions=G.InteractionParams(4.0, -0.3),
                                zeeman=G.ZeemanParams(0.1, 0.05), potential=G.HarmonicTrap(1.0),
                                sim_params=sp, psi_init=psi0)

    first = ws(40)
    run_simulation_checkpointed(first, checkpoint_dir=str(tmp_path), checkpoint_every=20)
    second = ws(100)
    run_simulation_checkpointed(second, checkpoint_dir=str(tmp_path), checkpoint_every=20,
                                resume=True)
    ows = O.make_workspace(og, 1, psi0, DT, False, omega=(1.0,), p=0.1, q=0.05, c0=4.0, c1=-0.3,
                           n_steps=100, save_every=20)
    O.run_simulation(ows)
    assert int(second.step[0]) == 100
    assert O.relative_l2(second.get_psi(0), ows.psi) <= 1e-10
    eo = O.total_energy(ows)
    assert abs(A.total_energy(second) - eo) <= 1e-12 * abs(eo)
