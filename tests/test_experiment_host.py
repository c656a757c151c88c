"""YAML experiment configuration (experiment.jl) on the host: the reference's parsing tests
(test_experiment.jl) restated, plus potential and Zeeman-ramp construction.  CPU only."""
import math

import numpy as np
import pytest

import becgpu as G
from becgpu import experiment as X


def test_interpolate_value():
    c = X.ConstantValue(3.0)
    assert [X.interpolate_value(c, t) for t in (0.0, 0.5, 1.0)] == [3.0, 3.0, 3.0]
    r = X.LinearRamp(1.0, 5.0)
    assert X.interpolate_value(r, 0.0) == 1.0
    assert X.interpolate_value(r, 0.5) == 3.0
    assert X.interpolate_value(r, 1.0) == 5.0
    assert X.interpolate_value(r, -0.1) == 1.0   # clamped
    assert X.interpolate_value(r, 1.5) == 5.0


FULL = """
experiment:
  name: "test experiment"
  system:
    atom: Rb87
    grid:
      n_points: [64]
      box_size: [20.0]
    interactions:
      c0: 10.0
      c1: -0.5
  ground_state:
    dt: 0.005
    n_steps: 100
    tol: 1.0e-8
    initial_state: polar
    zeeman:
      p: 0.0
      q: 0.1
    potential:
      type: harmonic
      omega: [1.0]
  sequence:
    - name: ramp
      duration: 1.0
      dt: 0.01
      save_every: 10
      zeeman:
        p:
          from: 0.0
          to: 0.5
        q: 0.0
      potential:
        type: harmonic
        omega: [1.0]
    - name: hold
      duration: 2.0
      dt: 0.01
      save_every: 20
      zeeman:
        p: 0.5
        q: 0.0
    - name: release
      duration: 0.5
      dt: 0.005
      save_every: 10
      zeeman:
        p: 0.0
        q: 0.0
      potential:
        type: none
"""


def test_yaml_parsing_full():
    c = X.load_experiment_from_string(FULL)
    assert c.name == "test experiment"
    assert c.system.atom_name == "Rb87"
    assert c.system.grid_n_points == [64] and c.system.grid_box_size == [20.0]
    assert c.system.interactions.c0 == 10.0 and c.system.interactions.c1 == -0.5
    assert c.system.ddi.enabled is False
    gs = c.ground_state
    assert (gs.dt, gs.n_steps, gs.tol, gs.initial_state) == (0.005, 100, 1e-8, "polar")
    assert (gs.zeeman.p, gs.zeeman.q) == (0.0, 0.1)
    assert gs.potential.type == "harmonic" and gs.potential.params["omega"] == [1.0]
    assert len(c.sequence) == 3
    ramp, hold, rel = c.sequence
    assert (ramp.name, ramp.duration, ramp.dt, ramp.save_every) == ("ramp", 1.0, 0.01, 10)
    assert isinstance(ramp.zeeman_p, X.LinearRamp) and (ramp.zeeman_p.from_, ramp.zeeman_p.to) == (0.0, 0.5)
    assert isinstance(ramp.zeeman_q, X.ConstantValue) and ramp.zeeman_q.value == 0.0
    assert ramp.potential is not None
    assert hold.name == "hold" and hold.potential is None        # inherits
    assert rel.name == "release" and rel.potential.type == "none"


def test_yaml_parsing_ddi_minimal_gravity_composite_noise():
    c = X.load_experiment_from_string("""
experiment:
  name: "ddi test"
  system:
    atom: Eu151
    grid: {n_points: [32], box_size: [10.0]}
    interactions: {c0: 5.0, c1: 0.0}
    ddi: {enabled: true, c_dd: 1.5e-5}
  sequence: []
""")
    assert c.system.ddi.enabled is True and c.system.ddi.c_dd == 1.5e-5
    assert c.system.atom_name == "Eu151"
    c = X.load_experiment_from_string("""
experiment:
  name: "minimal"
  system:
    atom: Na23
    grid: {n_points: 32, box_size: 10.0}
    interactions: {c0: 1.0, c1: 0.1}
  sequence: []
""")
    assert c.system.atom_name == "Na23" and c.system.grid_n_points == [32]
    assert c.ground_state is None and c.sequence == []
    c = X.load_experiment_from_string("""
experiment:
  name: "composite"
  system:
    atom: Rb87
    grid: {n_points: [32], box_size: [10.0]}
    interactions: {c0: 1.0, c1: 0.0}
  ground_state:
    dt: 0.005
    n_steps: 100
    tol: 1.0e-8
    potential:
      - {type: harmonic, omega: [1.0]}
      - {type: gravity, g: 9.81, axis: 1}
  sequence:
    - {name: noisy, duration: 1.0, dt: 0.01, noise_amplitude: 0.05, zeeman: {p: 0.0, q: 0.0},
       potential: {type: harmonic, omega: [1.0]}}
    - {name: quiet, duration: 1.0, dt: 0.01, zeeman: {p: 0.0, q: 0.0}}
""")
    pot = c.ground_state.potential
    assert pot.type == "composite"
    comps = pot.params["components"]
    assert [p.type for p in comps] == ["harmonic", "gravity"]
    assert comps[1].params == {"g": 9.81, "axis": 1}
    assert c.sequence[0].noise_amplitude == 0.05 and c.sequence[1].noise_amplitude is None


def test_crossed_dipole_parses_but_is_not_built():
    c = X.load_experiment_from_string("""
experiment:
  name: "dipole"
  system:
    atom: Rb87
    grid: {n_points: [8, 8, 8], box_size: [4.0, 4.0, 4.0]}
    interactions: {c0: 1.0, c1: 0.0}
  ground_state:
    dt: 0.005
    n_steps: 50
    tol: 1.0e-6
    potential:
      type: crossed_dipole
      polarizability: 1.5e-37
      beams:
        - {wavelength: 1064.0e-9, power: 10.0, waist: 50.0e-6, position: [0, 0, 0], direction: [1, 0, 0]}
        - {wavelength: 1064.0e-9, power: 10.0, waist: 50.0e-6, position: [0, 0, 0], direction: [0, 1, 0]}
  sequence: []
""")
    gs = c.ground_state
    assert gs.potential.type == "crossed_dipole"
    assert gs.potential.params["polarizability"] == 1.5e-37
    assert len(gs.potential.params["beams"]) == 2
    with pytest.raises(NotImplementedError):
        X.build_potential(gs.potential, 3)


def test_c_extra_sparse_and_constraint():
    """experiment.jl:233-249 (sparse cN keys fill zeros); interactions.jl:182-187 (c_total)."""
    c = X.load_experiment_from_string("""
experiment:
  system:
    atom: Eu151
    grid: {n_points: [16], box_size: [8.0]}
    interactions: {c_total: 12.0, c1_ratio: 0.1, c4: 0.3, c6: -0.2}
""")
    ip = c.system.interactions
    assert ip.c_extra == (0.0, 0.0, 0.3, 0.0, -0.2)
    assert ip.c0 == pytest.approx(12.0 / (1 + 36 * 0.1)) and ip.c1 == pytest.approx(0.1 * ip.c0)
    assert ip.c0 + 36 * ip.c1 == pytest.approx(12.0)
    with pytest.raises(ValueError):
        X.resolve_atom("K39")


def test_potentials_and_ramp_builders():
    g = G.make_grid(G.GridConfig((8, 4), (4.0, 2.0)))
    pot = X.build_potential(X.PotentialConfig("composite", {"components": [
        X.PotentialConfig("harmonic", {"omega": [1.0, 2.0]}),
        X.PotentialConfig("gravity", {"g": 0.5, "axis": 2})]}), 2)
    V = G.physics.evaluate_potential(pot, g)
    x, y = np.meshgrid(g.x[0], g.x[1])           # numpy C order (y, x)
    assert np.allclose(V, 0.5 * (x ** 2 + 4 * y ** 2) + 0.5 * y, atol=1e-14)
    with pytest.raises(ValueError):
        X.build_potential(X.PotentialConfig("harmonic", {"omega": [1.0]}), 2)
    with pytest.raises(ValueError):
        X.build_potential(X.PotentialConfig("gravity", {"axis": 3}), 2)
    ph = X._parse_phase({"name": "r", "duration": 2.0, "dt": 0.01,
                         "zeeman": {"p": {"from": 0.0, "to": 1.0}, "q": 0.3}})
    z = X.build_zeeman(ph, 5.0)
    assert isinstance(z, G.TimeDependentZeeman)
    assert G.zeeman_at(z, 6.0) == G.ZeemanParams(0.5, 0.3)
    assert G.zeeman_at(z, 4.0).p == 0.0 and G.zeeman_at(z, 9.0).p == 1.0   # clamped
    const = X._parse_phase({"name": "c", "duration": 1.0, "dt": 0.01, "zeeman": {"p": {"from": 0.2}}})
    assert X.build_zeeman(const, 0.0) == G.ZeemanParams(0.2, 0.0)
    ad = X._parse_phase({"name": "a", "duration": 1.0, "dt": 0.01, "adaptive_dt": {"tol": 1e-4}})
    assert (ad.adaptive_dt.dt_init, ad.adaptive_dt.dt_max, ad.adaptive_dt.tol) == (0.01, 0.1, 1e-4)
    assert math.isclose(ad.adaptive_dt.dt_min, 1e-5)
