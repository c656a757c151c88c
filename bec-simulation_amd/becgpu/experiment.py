"""YAML experiment runner over the GPU entry points (experiment.jl, experiment_runner.jl).

An experiment is an optional ground state followed by a sequence of real-time phases, each
with its own duration, dt, save cadence, Zeeman field (constants or linear ramps over the
phase), potential (inherited from the previous phase when absent), optional noise and
optional adaptive dt.  Parsing follows experiment.jl:89-249 key for key; running follows
run_experiment (experiment_runner.jl:54-155): find_ground_state, then per phase
make_workspace with the previous phase's psi, the time offset, noise, and
run_simulation! or run_simulation_adaptive!.

Differences from the reference, by necessity:
  * noise (_add_noise!, experiment_runner.jl:157-168) draws from this package's seeded
    SplitMix64 -> Box-Muller generator (physics.add_noise), not Julia's global RNG;
    run_experiment takes the seed (phase i uses seed + i);
  * `crossed_dipole` potentials (optical dipole traps of Gaussian beams) parse, but building
    one raises NotImplementedError: laser optics are not part of this package.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

import numpy as np
import yaml

from . import api as A
from .physics import (AdaptiveDtParams, CompositePotential, Eu151, GravityPotential, GridConfig,
                      HarmonicTrap, InteractionParams, LossParams, Na23, NoPotential, Rb87,
                      SimParams, TimeDependentZeeman, ZeemanParams, add_noise, cell_volume,
                      interaction_params_from_constraint, make_grid)


# --- ramps (experiment.jl:3-16) ---------------------------------------------------------
@dataclass(frozen=True)
class ConstantValue:
    value: float


@dataclass(frozen=True)
class LinearRamp:
    from_: float
    to: float


def interpolate_value(v, t_frac: float) -> float:
    """experiment.jl:15-16: constants, or from + (to - from) * clamp(t_frac, 0, 1)."""
    if isinstance(v, ConstantValue):
        return v.value
    return v.from_ + (v.to - v.from_) * min(max(t_frac, 0.0), 1.0)


# --- configuration (experiment.jl:18-75) ------------------------------------------------
@dataclass
class PotentialConfig:
    type: str
    params: dict = field(default_factory=dict)


@dataclass
class PhaseConfig:
    name: str
    duration: float
    dt: float
    save_every: int
    zeeman_p: object
    zeeman_q: object
    potential: PotentialConfig | None
    noise_amplitude: float | None
    adaptive_dt: AdaptiveDtParams | None


@dataclass
class GroundStateConfig:
    dt: float
    n_steps: int
    tol: float
    initial_state: str
    zeeman: ZeemanParams
    potential: PotentialConfig
    enable_ddi: bool


@dataclass
class DDIConfig:
    enabled: bool = False
    c_dd: float | None = None
    secular: bool = False


@dataclass
class SystemConfig:
    atom_name: str
    grid_n_points: list
    grid_box_size: list
    interactions: InteractionParams
    ddi: DDIConfig
    loss: LossParams | None


@dataclass
class ExperimentConfig:
    name: str
    system: SystemConfig
    ground_state: GroundStateConfig | None
    sequence: list


@dataclass
class ExperimentResult:
    config: ExperimentConfig
    ground_state_energy: float | None
    ground_state_converged: bool | None
    phase_results: list
    phase_names: list
    psi: np.ndarray | None = None     # final state (psi of the last phase or the ground state)


# --- atoms (experiment.jl:77-87) --------------------------------------------------------
ATOM_REGISTRY = {"Rb87": Rb87, "Na23": Na23, "Eu151": Eu151}


def resolve_atom(name: str):
    if name not in ATOM_REGISTRY:
        raise ValueError(f"Unknown atom: {name}. Available: {sorted(ATOM_REGISTRY)}")
    return ATOM_REGISTRY[name]


# --- YAML parsing (experiment.jl:89-249) ------------------------------------------------
def load_experiment(path: str) -> ExperimentConfig:
    with open(path) as f:
        return load_experiment_from_string(f.read())


def load_experiment_from_string(text: str) -> ExperimentConfig:
    data = yaml.safe_load(text)
    return _parse_experiment(data.get("experiment", data))


def _parse_experiment(d: dict) -> ExperimentConfig:
    name = d.get("name", "unnamed")
    system = _parse_system(d["system"])
    gs = _parse_ground_state(d["ground_state"]) if "ground_state" in d else None
    seq = [_parse_phase(p) for p in (d.get("sequence") or [])]
    return ExperimentConfig(name, system, gs, seq)


def _to_int_vec(v):
    return [int(x) for x in v] if isinstance(v, (list, tuple)) else [int(v)]


def _to_float_vec(v):
    return [float(x) for x in v] if isinstance(v, (list, tuple)) else [float(v)]


def _parse_c_extra(inter: dict):
    """experiment.jl:233-249: c2, c3, ... keys (sparse allowed) -> c_extra (c2 first)."""
    max_n = 0
    for k in inter:
        m = re.fullmatch(r"c(\d+)", str(k))
        if m and int(m.group(1)) >= 2:
            max_n = max(max_n, int(m.group(1)))
    if max_n < 2:
        return ()
    return tuple(float(inter.get(f"c{n}", 0.0)) for n in range(2, max_n + 1))


def _parse_system(d: dict) -> SystemConfig:
    atom_name = str(d["atom"])
    g = d["grid"]
    n_points = _to_int_vec(g["n_points"])
    box_size = _to_float_vec(g["box_size"])
    inter = d["interactions"]
    F_atom = resolve_atom(atom_name).F
    c_extra = _parse_c_extra(inter)
    if "c_total" in inter:
        interactions = interaction_params_from_constraint(
            c_total=float(inter["c_total"]), c1_ratio=float(inter.get("c1_ratio", 0.0)),
            F=F_atom, c_extra=c_extra)
    else:
        interactions = InteractionParams(float(inter["c0"]), float(inter["c1"]),
                                         float(inter.get("c_lhy", 0.0)), c_extra)
    ddi = DDIConfig()
    if "ddi" in d:
        dd = d["ddi"]
        ddi = DDIConfig(bool(dd.get("enabled", False)),
                        float(dd["c_dd"]) if "c_dd" in dd else None,
                        bool(dd.get("secular", False)))
    loss = None
    if "losses" in d:
        ld = d["losses"]
        loss = LossParams(float(ld.get("gamma_dr", 0.0)), float(ld.get("L3", 0.0)))
    return SystemConfig(atom_name, n_points, box_size, interactions, ddi, loss)


def _parse_ground_state(d: dict) -> GroundStateConfig:
    z = d.get("zeeman", {}) or {}
    return GroundStateConfig(
        dt=float(d["dt"]), n_steps=int(d["n_steps"]), tol=float(d["tol"]),
        initial_state=str(d.get("initial_state", "polar")),
        zeeman=ZeemanParams(float(z.get("p", 0.0)), float(z.get("q", 0.0))),
        potential=_parse_potential_config(d.get("potential", {"type": "none"})),
        enable_ddi=bool(d.get("enable_ddi", False)))


def _parse_ramp_or_constant(v):
    """experiment.jl:207-215: a number, {from, to} (a ramp) or {from} (a constant)."""
    if isinstance(v, dict):
        if "to" in v:
            return LinearRamp(float(v["from"]), float(v["to"]))
        return ConstantValue(float(v["from"]))
    return ConstantValue(float(v))


def _parse_phase(d: dict) -> PhaseConfig:
    dt = float(d["dt"])
    z = d.get("zeeman", {}) or {}
    adaptive = None
    if "adaptive_dt" in d:
        ad = d["adaptive_dt"] or {}
        adaptive = AdaptiveDtParams(dt_init=float(ad.get("dt_init", dt)),
                                    dt_min=float(ad.get("dt_min", 1e-5)),
                                    dt_max=float(ad.get("dt_max", 10 * dt)),
                                    tol=float(ad.get("tol", 1e-3)),
                                    error_mode=str(ad.get("error_mode", "step_change")))
    noise = d.get("noise_amplitude", None)
    return PhaseConfig(
        name=d.get("name", "unnamed"), duration=float(d["duration"]), dt=dt,
        save_every=int(d.get("save_every", 1)),
        zeeman_p=_parse_ramp_or_constant(z.get("p", 0.0)),
        zeeman_q=_parse_ramp_or_constant(z.get("q", 0.0)),
        potential=_parse_potential_config(d["potential"]) if "potential" in d else None,
        noise_amplitude=None if noise is None else float(noise),
        adaptive_dt=adaptive)


def _parse_potential_config(v) -> PotentialConfig:
    """experiment.jl:217-229: a dict {type, params...}, or a list (a composite)."""
    if isinstance(v, list):
        return PotentialConfig("composite", {"components": [_parse_potential_config(c) for c in v]})
    v = v or {}
    return PotentialConfig(str(v.get("type", "none")), {k: x for k, x in v.items() if k != "type"})


# --- running (experiment_runner.jl:1-155) -----------------------------------------------
def build_potential(pc: PotentialConfig, ndim: int):
    """_build_potential (experiment_runner.jl:1-27)."""
    t = pc.type
    if t == "none":
        return NoPotential()
    if t == "harmonic":
        omega = pc.params.get("omega")
        if omega is None:
            raise ValueError("Harmonic potential requires omega")
        omega = _to_float_vec(omega)
        if len(omega) != ndim:
            raise ValueError(f"omega length must match grid dimensions ({ndim})")
        return HarmonicTrap(*omega)
    if t == "gravity":
        axis = int(pc.params.get("axis", ndim))
        if not 1 <= axis <= ndim:
            raise ValueError(f"axis must be between 1 and {ndim}")
        return GravityPotential(float(pc.params.get("g", 9.81)), axis)
    if t == "composite":
        return CompositePotential(tuple(build_potential(c, ndim) for c in pc.params["components"]))
    if t == "crossed_dipole":
        raise NotImplementedError("crossed_dipole potentials (laser optics) are not built")
    raise ValueError(f"Unknown potential type: {t}")


def build_zeeman(phase: PhaseConfig, t_offset: float):
    """_build_zeeman (experiment_runner.jl:39-56): constants, or the ramp over the phase as a
    function of the absolute time."""
    if isinstance(phase.zeeman_p, ConstantValue) and isinstance(phase.zeeman_q, ConstantValue):
        return ZeemanParams(phase.zeeman_p.value, phase.zeeman_q.value)
    duration = phase.duration

    def field_at(t):
        frac = (t - t_offset) / duration if duration > 0 else 0.0
        return ZeemanParams(interpolate_value(phase.zeeman_p, frac),
                            interpolate_value(phase.zeeman_q, frac))
    return TimeDependentZeeman(field_at)


def run_experiment(config: ExperimentConfig, *, verbose=True, noise_seed=42, precision="fp64",
                   keep_snapshots=False) -> ExperimentResult:
    """run_experiment (experiment_runner.jl:58-155)."""
    sysc = config.system
    atom = resolve_atom(sysc.atom_name)
    ndim = len(sysc.grid_n_points)
    grid = make_grid(GridConfig(tuple(sysc.grid_n_points), tuple(sysc.grid_box_size)))
    c_dd = float("nan") if sysc.ddi.c_dd is None else sysc.ddi.c_dd

    gs_energy = gs_converged = None
    psi = None
    if config.ground_state is not None:
        gs = config.ground_state
        if verbose:
            print(f"Finding ground state ({gs.n_steps} steps, tol={gs.tol})...")
        r = A.find_ground_state(grid=grid, atom=atom, interactions=sysc.interactions,
                                zeeman=gs.zeeman, potential=build_potential(gs.potential, ndim),
                                dt=gs.dt, n_steps=gs.n_steps, tol=gs.tol,
                                initial_state=gs.initial_state, enable_ddi=gs.enable_ddi,
                                c_dd=c_dd, secular_ddi=sysc.ddi.secular, precision=precision)
        gs_energy, gs_converged = float(r["energy"]), bool(r["converged"])
        psi = r["workspace"].get_psi()
        r["workspace"].close()
        if verbose:
            print(f"  converged={gs_converged}, E={gs_energy}")

    results, names = [], []
    t_offset = 0.0
    prev_pot = config.ground_state.potential if config.ground_state is not None else None
    for i, phase in enumerate(config.sequence):
        if verbose:
            print(f"Phase {i + 1}: {phase.name} (duration={phase.duration}, dt={phase.dt})")
        pot_cfg = phase.potential if phase.potential is not None else prev_pot
        if pot_cfg is None:
            raise ValueError(f"Phase '{phase.name}' has no potential and nothing to inherit")
        prev_pot = pot_cfg
        n_steps = int(round(phase.duration / phase.dt))
        sp = SimParams(dt=phase.dt, n_steps=n_steps, save_every=phase.save_every)
        ws = A.make_workspace(grid=grid, atom=atom, interactions=sysc.interactions,
                              zeeman=build_zeeman(phase, t_offset),
                              potential=build_potential(pot_cfg, ndim), sim_params=sp,
                              psi_init=psi, enable_ddi=sysc.ddi.enabled, c_dd=c_dd,
                              secular_ddi=sysc.ddi.secular, loss=sysc.loss, precision=precision)
        ws.t[0] = t_offset
        if phase.noise_amplitude is not None and phase.noise_amplitude > 0:
            # _add_noise! (experiment_runner.jl:157-168): every non-dominant component, then
            # unit norm
            p0 = ws.get_psi()
            noisy = add_noise(p0, phase.noise_amplitude, noise_seed + i)
            noisy /= math.sqrt(float(np.sum(np.abs(noisy) ** 2)) * cell_volume(grid))
            ws.set_psi(noisy, 0)
        if phase.adaptive_dt is not None:
            out = A.run_simulation_adaptive(ws, adaptive=phase.adaptive_dt,
                                            t_end=t_offset + phase.duration,
                                            save_interval=phase.dt * phase.save_every,
                                            keep_snapshots=keep_snapshots)
            res = out["result"]
            if verbose:
                print(f"  adaptive: {out['n_accepted']} accepted, {out['n_rejected']} rejected, "
                      f"final_dt={out['final_dt']:.3g}")
        else:
            res = A.run_simulation(ws, keep_snapshots=keep_snapshots)
        psi = ws.get_psi()
        t_offset += phase.duration
        results.append(res)
        names.append(phase.name)
        if verbose:
            print(f"  final t={ws.t[0]}, E={res.energies[-1]}")
        ws.close()
    return ExperimentResult(config, gs_energy, gs_converged, results, names, psi)
