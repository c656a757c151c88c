"""YAML experiment runner on the GPU (experiment_runner.jl:54-155).

The reference's run_experiment integration tests (test_experiment.jl) restated, and parity of
real-time phases with a time-dependent Zeeman ramp against the FP64 oracle given the same
field function: the leapfrog of run_simulation! (every half potential step reads the field
at the current t, simulation.jl:49-94, split_step.jl:49-50) with loss, and
run_simulation_adaptive! in both error modes.  Bar: psi relative L2 <= 1e-10, energies
<= 1e-12 relative.
"""
import math

import numpy as np
import pytest

from oracle import spinor_oracle as O
import becgpu as G
from becgpu import experiment as X
from helpers import grids, pair, random_psi, rel_l2

pytestmark = pytest.mark.gpu

BASE = """
experiment:
  name: "{name}"
  system:
    atom: Rb87
    grid: {{n_points: [32], box_size: [20.0]}}
    interactions: {{c0: 10.0, c1: -0.5}}
  ground_state:
    dt: 0.005
    n_steps: 200
    tol: 1.0e-6
    initial_state: polar
    zeeman: {{p: 0.0, q: 0.1}}
    potential:
{gs_pot}
  sequence:
    - name: {phase}
      duration: 0.1
      dt: 0.001
      save_every: 50
{extra}
      zeeman:
        p: {p}
        q: 0.1
      potential: {{type: harmonic, omega: [1.0]}}
"""
HARMONIC = "      type: harmonic\n      omega: [1.0]"
COMPOSITE = ("      - {type: harmonic, omega: [1.0]}\n"
             "      - {type: gravity, g: 0.1, axis: 1}")


@pytest.mark.parametrize("case", ["ramp", "noise", "composite"])
def test_run_experiment_integration(case):
    text = BASE.format(name=case, gs_pot=COMPOSITE if case == "composite" else HARMONIC,
                       phase="evolve", extra="      noise_amplitude: 0.01" if case == "noise" else "",
                       p="{from: 0.0, to: 0.1}" if case == "ramp" else "0.0")
    cfg = X.load_experiment_from_string(text)
    r = X.run_experiment(cfg, verbose=False)
    assert isinstance(r.ground_state_energy, float) and math.isfinite(r.ground_state_energy)
    assert isinstance(r.ground_state_converged, bool)
    assert len(r.phase_results) == 1 and r.phase_names == ["evolve"]
    sim = r.phase_results[0]
    assert len(sim.times) > 1 and sim.times[0] == 0.0
    assert all(n > 0 for n in sim.norms)
    assert sim.times[-1] == pytest.approx(0.1, abs=1e-12)
    # the ground state the phase starts from: oracle ITP with the same settings
    og, _ = grids((32,), (20.0,))
    V = O.harmonic_potential(og, (1.0,))
    if case == "composite":
        V = V + 0.1 * np.asarray(og.x[0])
    ows = O.make_workspace(og, 1, O.init_psi(og, 1, "polar"), 0.005, True, V=V, q=0.1, c0=10.0,
                           c1=-0.5, n_steps=200)
    ro = O.find_ground_state(ows, 200, 1e-6)
    assert abs(r.ground_state_energy - ro["energy"]) <= 1e-12 * abs(ro["energy"])
    assert r.ground_state_converged == ro["converged"]


def _ramp(t0, dur):
    def f(t):
        s = min(max((t - t0) / dur, 0.0), 1.0)
        return 0.5 * s, 0.1 - 0.3 * s
    return f


def test_time_dependent_zeeman_leapfrog_with_loss():
    n, L = (32, 32), (12.0, 10.0)
    og, _ = grids(n, L)
    psi = random_psi(og.shape, 3, 5) * O.gaussian(og, (2.5, 2.0))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    t0, dt, steps = 0.3, 0.002, 40
    ows, gws = pair(n, L, 1, psi, dt, False, omega=(1, 1), c0=10.0, c1=-0.5, n_steps=steps,
                    save_every=10, loss=(0.05, 0.01))
    fo = _ramp(t0, steps * dt)
    ows.zeeman_fn = fo
    ows.t = t0
    gws.zeeman_fn = lambda t: G.ZeemanParams(*fo(t))
    gws.t[0] = t0
    ro = O.run_simulation(ows)
    rg = G.run_simulation(gws, keep_snapshots=False)
    assert np.allclose(rg.times, ro["times"], rtol=0, atol=1e-12)
    for a, b in zip(rg.energies, ro["energies"]):
        assert abs(a - b) <= 1e-12 * abs(b), (a, b)
    assert rel_l2(gws.get_psi(), ows.psi) <= 1e-10
    # the field did change: a constant field at the start value gives another state
    _, gc = pair(n, L, 1, psi, dt, False, omega=(1, 1), p=0.0, q=0.1, c0=10.0, c1=-0.5,
                 n_steps=steps, save_every=10, loss=(0.05, 0.01))
    G.run_simulation(gc, keep_snapshots=False)
    assert rel_l2(gc.get_psi(), ows.psi) > 1e-6


@pytest.mark.parametrize("mode", ["step_change", "richardson"])
def test_time_dependent_zeeman_adaptive(mode):
    n, L = (32, 32), (12.0, 10.0)
    og, _ = grids(n, L)
    psi = O.init_psi(og, 1, "uniform") * (1 + 0.05 * np.random.default_rng(3).normal(size=og.shape))[None]
    psi = psi + 0.02 * random_psi(og.shape, 3, 4)
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 1, psi, 0.001, False, omega=(1, 1), c0=10.0, c1=-0.5)
    fo = _ramp(0.0, 0.15)
    ows.zeeman_fn = fo
    gws.zeeman_fn = lambda t: G.ZeemanParams(*fo(t))
    ro = O.run_simulation_adaptive(ows, dt_init=0.002, dt_min=1e-5, dt_max=0.01, tol=1e-4,
                                   error_mode=mode, t_end=0.15, save_interval=0.05)
    rg = G.run_simulation_adaptive(gws, adaptive=G.AdaptiveDtParams(dt_init=0.002, dt_min=1e-5,
                                                                    dt_max=0.01, tol=1e-4,
                                                                    error_mode=mode),
                                   t_end=0.15, save_interval=0.05)
    assert (rg["n_accepted"], rg["n_rejected"]) == (ro["n_accepted"], ro["n_rejected"])
    for a, b in zip(rg["result"].energies, ro["result"]["energies"]):
        assert abs(a - b) <= 1e-12 * abs(b)
    assert rel_l2(gws.get_psi(), ows.psi) <= 1e-10
