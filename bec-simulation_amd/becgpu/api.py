"""Public entry points mirroring SpinorBEC.jl, running the hot path on the GPU.

make_workspace (initialization.jl:86-191), split_step! (split_step.jl:10-35),
run_simulation! (simulation.jl:1-94), find_ground_state (ground_state.jl:80-163),
find_ground_state_multistart (ground_state.jl:239-271), scan_continuation
(ground_state.jl:328-378), run_simulation_yoshida! (yoshida.jl:11-98),
total_energy (energy.jl:6-54), total_norm / magnetization (observables.jl:1-33).

Julia's `!` suffix is dropped (split_step, run_simulation, ...); the functions
mutate the workspace exactly as the reference's do.  psi arrays are numpy
complex128 of shape (D, nz, ny, nx) -- byte-identical to the reference's
column-major (nx, ny, nz, D) array.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .physics import (AdaptiveDtParams, Grid, HarmonicTrap, InteractionParams, NoPotential,
                      SimParams, SpinSystem, ZeemanParams, c0c1_to_gS, c_extra_to_delta_gS,
                      cell_volume, classify_phase, dipolar_relaxation_rates, evaluate_potential,
                      init_psi, spin_matrices, zeeman_energies, TimeDependentZeeman,
                      zeeman_at)

ITP_EXPONENT_LIMIT = 50.0          # ground_state.jl:1


@dataclass
class SimulationResult:
    """simulation_utils.jl:1-7."""
    times: list = field(default_factory=list)
    energies: list = field(default_factory=list)
    norms: list = field(default_factory=list)
    magnetizations: list = field(default_factory=list)
    psi_snapshots: list = field(default_factory=list)


def _precision_code(precision):
    p = str(precision).lower()
    if p in ("fp64", "float64", "complexf64", "double"):
        return L.BEC_FP64
    if p in ("fp32", "float32", "complexf32", "fp32-bulk", "fp32_bulk"):
        return L.BEC_FP32
    raise ValueError(f"unknown precision {precision!r}")


class Workspace:
    """Device-resident workspace of B instances of one shape (types.jl:387-407)."""

    def __init__(self, *, grid: Grid, F: int, n_instances: int, sim_params: SimParams,
                 precision="fp64", enable_ddi=False, secular_ddi=False, ddi_padding=False,
                 tensor=False, autotune=True):
        self.grid = grid
        self.F = F
        self.sys = SpinSystem(F)
        self.D = 2 * F + 1
        self.B = n_instances
        self.sim_params = sim_params
        self.precision = _precision_code(precision)
        self.enable_ddi = bool(enable_ddi)
        self.tensor = bool(tensor)
        cfg = L.BecConfig()
        nd = grid.ndim
        cfg.ndim = nd
        for d in range(3):
            cfg.n[d] = grid.config.n_points[d] if d < nd else 1
            cfg.L[d] = grid.config.box_size[d] if d < nd else 1.0
        cfg.F = F
        cfg.n_instances = n_instances
        cfg.precision = self.precision
        cfg.imaginary_time = 1 if sim_params.imaginary_time else 0
        cfg.normalize_every = int(sim_params.normalize_every)
        cfg.enable_ddi = 1 if enable_ddi else 0
        cfg.ddi_secular = 1 if secular_ddi else 0
        cfg.ddi_padding = 1 if ddi_padding else 0
        cfg.tensor = 1 if tensor else 0
        xt = np.ascontiguousarray(np.concatenate(grid.x), dtype=np.float64)
        kt = np.ascontiguousarray(np.concatenate(grid.k), dtype=np.float64)
        self._keep = (xt, kt)
        cfg.x_tab = xt.ctypes.data_as(C.POINTER(C.c_double))
        cfg.k_tab = kt.ctypes.data_as(C.POINTER(C.c_double))
        self._cfg = cfg
        h = C.c_void_p()
        L.check(L.lib().bec_ws_create(C.byref(cfg), C.byref(h)), "bec_ws_create")
        self._h = h
        self.t = np.zeros(n_instances)
        self.step = np.zeros(n_instances, dtype=np.int64)
        self.params = [None] * n_instances
        self.interactions = [None] * n_instances
        self.zeeman = [ZeemanParams()] * n_instances
        # TimeDependentZeeman.B_func (t -> ZeemanParams) shared by the instances, or None;
        # the drivers apply it at the instance's time before every half potential step and
        # energy, as zeeman_at(ws.zeeman, ws.state.t) does (split_step.jl:49-50, energy.jl:15)
        self.zeeman_fn = None
        self._loss = [False] * n_instances
        self._active = list(range(n_instances))
        self.xrows = self.autotune() if autotune else None

    # -- handle ------------------------------------------------------------
    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("workspace destroyed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            L.lib().bec_ws_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def bytes_for(grid: Grid, F: int, n_instances=1, precision="fp64", enable_ddi=False):
        cfg = L.BecConfig()
        nd = grid.ndim
        cfg.ndim = nd
        for d in range(3):
            cfg.n[d] = grid.config.n_points[d] if d < nd else 1
        cfg.F, cfg.n_instances = F, n_instances
        cfg.precision = _precision_code(precision)
        cfg.enable_ddi = 1 if enable_ddi else 0
        return int(L.lib().bec_workspace_bytes(C.byref(cfg)))

    # -- parameters / state --------------------------------------------------
    def set_instance(self, inst, *, interactions: InteractionParams, zeeman=ZeemanParams(),
                     potential=NoPotential(), dt=None, c_dd=0.0):
        p = L.BecInstParams()
        p.dt = float(self.sim_params.dt if dt is None else dt)
        p.c0, p.c1, p.c_lhy = interactions.c0, interactions.c1, interactions.c_lhy
        # get_cn(ip, 2) (types.jl:157-162): the nematic coupling is c_extra[1]
        p.c2 = float(interactions.c_extra[0]) if len(interactions.c_extra) >= 1 else 0.0
        p.c_dd = float(c_dd)
        p.p, p.q = zeeman.p, zeeman.q
        if isinstance(potential, HarmonicTrap):
            for d in range(self.grid.ndim):
                p.omega[d] = potential.omega[d]
            L.check(L.lib().bec_set_params(self.handle, inst, C.byref(p)), "bec_set_params")
        elif isinstance(potential, NoPotential) or potential is None:
            L.check(L.lib().bec_set_params(self.handle, inst, C.byref(p)), "bec_set_params")
        else:
            L.check(L.lib().bec_set_params(self.handle, inst, C.byref(p)), "bec_set_params")
            V = np.ascontiguousarray(evaluate_potential(potential, self.grid), dtype=np.float64)
            L.check(L.lib().bec_set_potential(self.handle, inst, V.ctypes.data), "bec_set_potential")
        self.params[inst] = p
        self.interactions[inst] = interactions
        self.zeeman[inst] = zeeman

    def sync_zeeman(self):
        """Set each active instance's Zeeman field to zeeman_fn(t) (no-op without one)."""
        if self.zeeman_fn is None:
            return
        for inst in self._active:
            z = self.zeeman_fn(float(self.t[inst]))
            if z == self.zeeman[inst] or self.params[inst] is None:
                continue
            p = self.params[inst]
            p.p, p.q = float(z.p), float(z.q)
            L.check(L.lib().bec_set_params(self.handle, inst, C.byref(p)), "bec_set_params")
            self.zeeman[inst] = z

    def set_loss(self, inst, loss):
        """LossParams(gamma_dr, L3) of one instance, or None (types.jl:291-296). The
        per-component rates are formed here (losses.jl:60-82) and applied on the device."""
        if loss is None or (loss.gamma_dr < 1e-30 and loss.L3 < 1e-30):   # losses.jl:31
            L.check(L.lib().bec_set_loss(self.handle, inst, None, 0), "bec_set_loss")
            self._loss[inst] = False
            return
        rates = np.ascontiguousarray(
            [g + loss.L3 for g in dipolar_relaxation_rates(self.F, loss.gamma_dr)], dtype=np.float64)
        L.check(L.lib().bec_set_loss(self.handle, inst, rates.ctypes.data, self.D), "bec_set_loss")
        self._loss[inst] = True

    def set_raman(self, inst, raman):
        """RamanCoupling(Omega_R, delta, k_eff) of one instance, or None (types.jl:179-183):
        apply_raman_step! (raman.jl:8-31) runs on the device in every half potential step."""
        if raman is None:
            L.check(L.lib().bec_set_raman(self.handle, inst, None), "bec_set_raman")
            return
        k = tuple(float(v) for v in raman.k_eff)
        if len(k) != self.grid.ndim:
            raise ValueError(f"RamanCoupling.k_eff has {len(k)} entries for a {self.grid.ndim}-D grid")
        r = np.ascontiguousarray([float(raman.Omega_R), float(raman.delta), *k], dtype=np.float64)
        L.check(L.lib().bec_set_raman(self.handle, inst, r.ctypes.data), "bec_set_raman")

    @property
    def has_loss(self):
        return any(self._loss[i] for i in self.active)

    def set_dt(self, inst, dt):
        """Change one instance's step (the dt of _rebuild_workspace_with_dt,
        ground_state.jl:380-392); every other parameter is kept."""
        p = self.params[inst]
        p.dt = float(dt)
        L.check(L.lib().bec_set_params(self.handle, inst, C.byref(p)), "bec_set_params")

    def normalize_constrained(self, target_mz):
        """_normalize_psi_constrained! (ground_state.jl:279-320) of the active instances;
        target_mz: one value for all, or a sequence of B values."""
        t = np.full(self.B, float(target_mz)) if np.isscalar(target_mz) else \
            np.ascontiguousarray(target_mz, dtype=np.float64)
        L.check(L.lib().bec_normalize_constrained(self.handle, t.ctypes.data),
                "bec_normalize_constrained")

    def set_tensor(self, inst, channels: dict):
        """g_S couplings {S: g} of the general-F tensor step for one instance."""
        g = np.array([channels.get(S, 0.0) for S in range(0, 2 * self.F + 1, 2)], dtype=np.float64)
        L.check(L.lib().bec_set_tensor(self.handle, inst, g.ctypes.data, len(g)), "bec_set_tensor")

    def set_psi(self, psi, inst=0):
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        if a.shape != (self.D,) + self.grid.shape:
            raise ValueError(f"psi shape {a.shape} != {(self.D,) + self.grid.shape}")
        L.check(L.lib().bec_set_psi(self.handle, inst, a.ctypes.data), "bec_set_psi")

    def get_psi(self, inst=0):
        out = np.empty((self.D,) + self.grid.shape, dtype=np.complex128)
        L.check(L.lib().bec_get_psi(self.handle, inst, out.ctypes.data), "bec_get_psi")
        return out

    @property
    def psi(self):
        return self.get_psi(0)

    def set_active(self, insts):
        a = np.ascontiguousarray(list(insts), dtype=np.int32)
        L.check(L.lib().bec_set_active(self.handle, a.ctypes.data, len(a)), "bec_set_active")
        self._active = [int(v) for v in a]

    @property
    def active(self):
        return list(self._active)

    # -- device ops ------------------------------------------------------------
    def energy_terms(self):
        out = np.zeros((self.B, L.E_COUNT))
        L.check(L.lib().bec_energy(self.handle, out.ctypes.data), "bec_energy")
        return out

    def diff(self, a, b, mode):
        out = np.zeros((self.B, 2))
        L.check(L.lib().bec_diff(self.handle, a, b, mode, out.ctypes.data), "bec_diff")
        return out

    def copy_slot(self, dst, src):
        L.check(L.lib().bec_copy_slot(self.handle, dst, src), "bec_copy_slot")

    def swap_slots(self, a, b):
        """Exchange two slots for every instance (bec_swap_slots: no data moves)."""
        L.check(L.lib().bec_swap_slots(self.handle, a, b), "bec_swap_slots")

    def itp_chunk(self, n_steps):
        """bec_itp_chunk: n_steps ITP split steps, then per instance
        [energy, dpsi vs slot 1, norm, max |Phi|, status] and slot 1 <- psi."""
        out = np.zeros((self.B, L.CHUNK_COUNT))
        L.check(L.lib().bec_itp_chunk(self.handle, int(n_steps), out.ctypes.data),
                "bec_itp_chunk")
        return out

    def spin_density(self, inst=0):
        """_compute_spin_density! (observables.jl:61-88): (Fx, Fy, Fz), each of the grid's
        numpy shape, in FP64."""
        out = np.zeros((3,) + self.grid.shape)
        L.check(L.lib().bec_spin_density(self.handle, int(inst), out.ctypes.data),
                "bec_spin_density")
        return out[0], out[1], out[2]

    def itp_guard(self):
        out = np.zeros((self.B, 2))
        L.check(L.lib().bec_itp_guard(self.handle, out.ctypes.data), "bec_itp_guard")
        return out

    def profile(self, on: bool):
        L.check(L.lib().bec_profile_enable(self.handle, 1 if on else 0), "bec_profile_enable")

    def profile_read(self):
        """{kernel class: (total_ms, launches)} since profile(True)."""
        ms = np.zeros(len(L.K_NAMES))
        cnt = np.zeros(len(L.K_NAMES), dtype=np.int64)
        L.check(L.lib().bec_profile_read(self.handle, ms.ctypes.data, cnt.ctypes.data),
                "bec_profile_read")
        return {n: (float(ms[i]), int(cnt[i])) for i, n in enumerate(L.K_NAMES)}

    def profile_work(self):
        """Work since profile(True) or the last call: per-site operator counts
        {"tensor_matvecs", "rotations", "flops_per_matvec", "flops_per_rotation", "flops"},
        "fft_flops" of the fused x/site launches, and "bytes": algorithmic HBM bytes issued
        per kernel class."""
        K = len(L.K_NAMES)
        out = np.zeros(7 + K)
        L.check(L.lib().bec_profile_work(self.handle, out.ctypes.data), "bec_profile_work")
        tensor = out[0] * out[2] - out[5 + K] * (out[2] - out[6 + K])
        return {"tensor_matvecs": out[0], "rotations": out[1], "flops_per_matvec": out[2],
                "first_term_matvecs": out[5 + K], "flops_per_first_term": out[6 + K],
                "flops_per_rotation": out[3], "flops": tensor + out[1] * out[3],
                "fft_flops": float(out[4 + len(L.K_NAMES)]),
                "bytes": {n: float(out[4 + i]) for i, n in enumerate(L.K_NAMES)}}

    def set_xsite(self, on: bool) -> bool:
        """Run the heavy per-site operators of F = 2 and 6 in the fused x/site kernel (on,
        the default) or in separate x and lane-per-site passes; returns whether the fused
        kernel is in use for this workspace's shape."""
        r = L.lib().bec_set_xsite(self.handle, 1 if on else 0)
        L.check(min(r, 0), "bec_set_xsite")
        return r == 1

    def synchronize(self):
        L.check(L.lib().bec_synchronize(self.handle), "bec_synchronize")

    def autotune(self):
        """Pick the x-pass tile for this shape (timed once per shape and process; psi is not
        modified).  Returns the x rows per block in use."""
        r = C.c_int(0)
        L.check(L.lib().bec_autotune(self.handle, C.byref(r)), "bec_autotune")
        return int(r.value)

    def query(self):
        """True when all work queued on this workspace has finished (does not block)."""
        rc = L.lib().bec_query(self.handle)
        if rc < 0:
            L.check(rc, "bec_query")
        return rc == 0


# ---------------------------------------------------------------------------
# Spectral derivatives (fft_utils.jl)
# ---------------------------------------------------------------------------
def _grid_config(grid: Grid):
    cfg = L.BecConfig()
    nd = grid.ndim
    cfg.ndim = nd
    for d in range(3):
        cfg.n[d] = grid.config.n_points[d] if d < nd else 1
        cfg.L[d] = grid.config.box_size[d] if d < nd else 1.0
    cfg.F, cfg.n_instances = 1, 1
    xt = np.ascontiguousarray(np.concatenate(grid.x), dtype=np.float64)
    kt = np.ascontiguousarray(np.concatenate(grid.k), dtype=np.float64)
    cfg.x_tab = xt.ctypes.data_as(C.POINTER(C.c_double))
    cfg.k_tab = kt.ctypes.data_as(C.POINTER(C.c_double))
    return cfg, (xt, kt)


def fft_partial_derivative(f, grid: Grid, dim: int):
    """_fft_partial_derivative (fft_utils.jl:4-19): Re IFFT(i k_dim FFT(f)) of a real field
    of the grid's shape; dim is the reference's 1-based axis (1 = x, the fastest)."""
    a = np.ascontiguousarray(f, dtype=np.float64)
    if a.shape != grid.shape:
        raise ValueError(f"field shape {a.shape} != grid shape {grid.shape}")
    if not 1 <= dim <= grid.ndim:
        raise ValueError(f"dim must be in 1..{grid.ndim}")
    cfg, keep = _grid_config(grid)
    out = np.empty(grid.shape)
    L.check(L.lib().bec_fft_derivative(C.byref(cfg), a.ctypes.data, dim - 1, out.ctypes.data),
            "bec_fft_derivative")
    return out


def fft_gradient(f, grid: Grid):
    """_fft_gradient (fft_utils.jl:25-44): the tuple (d/dx, d/dy, ...) of a real field."""
    a = np.ascontiguousarray(f, dtype=np.float64)
    if a.shape != grid.shape:
        raise ValueError(f"field shape {a.shape} != grid shape {grid.shape}")
    cfg, keep = _grid_config(grid)
    out = np.empty((grid.ndim,) + grid.shape)
    L.check(L.lib().bec_fft_gradient(C.byref(cfg), a.ctypes.data, out.ctypes.data),
            "bec_fft_gradient")
    return tuple(out[d] for d in range(grid.ndim))


# ---------------------------------------------------------------------------
# make_workspace
# ---------------------------------------------------------------------------
def _tensor_active(F, interactions: InteractionParams):
    """initialization.jl:164-185: even-rank c_extra with 4 <= k <= 2F activates the
    general-F tensor step."""
    return any((i + 1) % 2 == 0 and 4 <= i + 1 <= 2 * F and abs(v) > 1e-30
               for i, v in enumerate(interactions.c_extra, start=1))


def tensor_channels(F, interactions: InteractionParams):
    """g_S = _c0c1_to_gS(F, c0, c1) + _c_extra_to_delta_gS(F, c_extra), channels with
    |g_S| <= 1e-30 dropped (initialization.jl:169-174, tensor_interaction.jl:106-119)."""
    gb = c0c1_to_gS(F, interactions.c0, interactions.c1)
    gd = c_extra_to_delta_gS(F, interactions.c_extra)
    g = {S: gb.get(S, 0.0) + gd.get(S, 0.0) for S in range(0, 2 * F + 1, 2)}
    return {S: v for S, v in g.items() if abs(v) > 1e-30}


def make_workspace(*, grid: Grid, atom, interactions: InteractionParams,
                   zeeman=ZeemanParams(), potential=NoPotential(), sim_params: SimParams,
                   psi_init=None, enable_ddi=False, c_dd=float("nan"), secular_ddi=False,
                   raman=None, loss=None, ddi_padding=False, precision="fp64", fft_flags=None):
    F = atom.F
    zeeman_fn = None
    if isinstance(zeeman, TimeDependentZeeman):    # types.jl:173-175
        zeeman_fn = zeeman.B_func
        zeeman = zeeman_at(zeeman, 0.0)
    tensor = _tensor_active(F, interactions)
    channels = None
    if tensor:
        channels = tensor_channels(F, interactions)
        # the tensor step carries every contact channel; c0 = c1 = 0 elsewhere
        interactions = InteractionParams(0.0, 0.0, interactions.c_lhy, ())
    c_dd_val = 0.0
    if enable_ddi:
        if math.isnan(c_dd):
            if atom.mu_mag > 0.0:
                raise ValueError(
                    f"enable_ddi=true for dipolar atom {atom.name} but c_dd not specified.")
            c_dd_val = 0.0
        else:
            c_dd_val = float(c_dd)
    ws = Workspace(grid=grid, F=F, n_instances=1, sim_params=sim_params, precision=precision,
                   enable_ddi=enable_ddi, secular_ddi=secular_ddi, ddi_padding=ddi_padding,
                   tensor=tensor)
    ws.set_instance(0, interactions=interactions, zeeman=zeeman, potential=potential,
                    c_dd=c_dd_val)
    if tensor:
        ws.set_tensor(0, channels)
    if loss is not None:
        ws.set_loss(0, loss)
    if raman is not None:
        ws.set_raman(0, raman)
    ws.zeeman_fn = zeeman_fn
    psi = init_psi(grid, SpinSystem(F)) if psi_init is None else np.array(psi_init, dtype=complex)
    ws.set_psi(psi, 0)
    ws.atom = atom
    ws.potential = potential
    return ws


# ---------------------------------------------------------------------------
# Observables
# ---------------------------------------------------------------------------
def energy_terms(ws: Workspace, inst=0):
    row = ws.energy_terms()[inst]
    return {name: float(row[i]) for i, name in enumerate(L.E_NAMES)}


def total_energy(ws: Workspace, inst=0):
    return float(ws.energy_terms()[inst, L.E_TOTAL])


def total_norm(ws: Workspace, inst=0):
    return float(ws.energy_terms()[inst, L.E_NORM])


def magnetization(ws: Workspace, inst=0):
    return float(ws.energy_terms()[inst, L.E_MZ])


def _record_snapshot(ws, res: SimulationResult, keep_psi=True, inst=0):
    ws.sync_zeeman()
    e = ws.energy_terms()[inst]
    res.times.append(float(ws.t[inst]))
    res.energies.append(float(e[L.E_TOTAL]))
    res.norms.append(float(e[L.E_NORM]))
    res.magnetizations.append(float(e[L.E_MZ]))
    if keep_psi:
        res.psi_snapshots.append(ws.get_psi(inst))


# ---------------------------------------------------------------------------
# Integrators
# ---------------------------------------------------------------------------
def split_step(ws: Workspace, n: int = 1, graph: bool = False):
    """split_step! (split_step.jl:10-35), n times.  graph: issue the n steps as one captured
    HIP graph (cached per step count and active set; same results)."""
    if ws.zeeman_fn is not None and n > 1 and not ws.sim_params.imaginary_time:
        for _ in range(n):            # the field follows t from step to step
            split_step(ws, 1, graph=False)
        return
    ws.sync_zeeman()
    if graph:
        L.check(L.lib().bec_split_step_graph(ws.handle, int(n)), "bec_split_step_graph")
    else:
        L.check(L.lib().bec_split_step(ws.handle, int(n)), "bec_split_step")
    if not ws.sim_params.imaginary_time:
        for i in ws.active:
            for _ in range(n):
                ws.t[i] += ws.params[i].dt
    ws.step[ws.active] += n


def run_simulation(ws: Workspace, callback=None, keep_snapshots=True):
    """run_simulation! (simulation.jl:1-94): ITP -> split_step loop with saves every
    save_every; real time -> leapfrog with merged V(dt) and splits at save points."""
    sp = ws.sim_params
    res = SimulationResult()
    _record_snapshot(ws, res, keep_snapshots)
    if sp.imaginary_time:
        done = 0
        while done < sp.n_steps:
            k = min(sp.save_every - done % sp.save_every, sp.n_steps - done)
            split_step(ws, k)
            done += k
            if done % sp.save_every == 0:
                _record_snapshot(ws, res, keep_snapshots)
                if callback is not None:
                    callback(ws, done)
        return res
    dt = ws.params[0].dt
    if ws.zeeman_fn is not None:
        return _run_leapfrog_time_dependent(ws, res, callback, keep_snapshots)
    done = 0
    while done < sp.n_steps:
        k = min(sp.save_every - done % sp.save_every, sp.n_steps - done)
        L.check(L.lib().bec_leapfrog_segment(ws.handle, k), "bec_leapfrog_segment")
        for _ in range(k):
            ws.t[0] += dt
        ws.step[0] += k
        done += k
        if done % sp.save_every == 0:
            _record_snapshot(ws, res, keep_snapshots)
            if callback is not None:
                callback(ws, done)
    return res


def _run_leapfrog_time_dependent(ws, res, callback, keep_snapshots):
    """_run_simulation_leapfrog! (simulation.jl:49-94) with a time-dependent Zeeman field:
    every half potential step reads the field at the current t (the merged V(dt) of a step
    at its start time), so the steps run as separate half potential and kinetic launches."""
    lib = L.lib()
    sp = ws.sim_params
    dt = ws.params[0].dt

    def hp(dth):
        ws.sync_zeeman()
        L.check(lib.bec_half_potential(ws.handle, float(dth), 0), "bec_half_potential")

    hp(dt / 2)
    for step in range(1, sp.n_steps + 1):
        L.check(lib.bec_kinetic(ws.handle, float(dt), 0), "bec_kinetic")
        is_save = step % sp.save_every == 0
        is_last = step == sp.n_steps
        split = is_save or is_last
        hp(dt / 2 if split else dt)
        if ws.has_loss:
            L.check(lib.bec_loss_step(ws.handle, float(dt)), "bec_loss_step")
        ws.t[0] += dt
        ws.step[0] += 1
        if is_save:
            _record_snapshot(ws, res, keep_snapshots)
            if callback is not None:
                callback(ws, step)
        if split and not is_last:
            hp(dt / 2)
    return res


def strang_core(ws: Workspace, dt: float, src_slot: int = 0):
    """_strang_core! (split_step.jl:146-153); src_slot > 0 starts from that slot's psi
    (kept) and leaves the result in slot 0."""
    ws.sync_zeeman()
    if src_slot:
        L.check(L.lib().bec_strang_core_from(ws.handle, float(dt), int(src_slot)),
                "bec_strang_core_from")
    else:
        L.check(L.lib().bec_strang_core(ws.handle, float(dt)), "bec_strang_core")


def yoshida_core(ws: Workspace, dt: float, src_slot: int = 0):
    """_yoshida_core! (split_step.jl:155-183); src_slot as strang_core."""
    ws.sync_zeeman()
    if src_slot:
        L.check(L.lib().bec_yoshida_core_from(ws.handle, float(dt), int(src_slot)),
                "bec_yoshida_core_from")
    else:
        L.check(L.lib().bec_yoshida_core(ws.handle, float(dt)), "bec_yoshida_core")


def run_simulation_yoshida(ws: Workspace, *, adaptive=AdaptiveDtParams(), t_end: float,
                           save_interval: float, callback=None, keep_snapshots=True,
                           max_accepted=None):
    """run_simulation_yoshida! (yoshida.jl:11-98): adaptive S4 with the embedded S2
    error estimate ||psi_Y - psi_S||^2/||psi_S||^2 and the reference's PI rules.
    psi_old, psi_strang stay on the device (slots 2 and 3)."""
    dt = min(max(adaptive.dt_init, adaptive.dt_min), adaptive.dt_max)
    res = SimulationResult()
    _record_snapshot(ws, res, keep_snapshots)
    next_save = ws.t[0] + save_interval
    n_acc = n_rej = 0
    while ws.t[0] < t_end - 1e-14:
        if max_accepted is not None and n_acc >= max_accepted:
            break
        t = ws.t[0]
        dt_step = min(dt, t_end - t)
        rem = next_save - t
        if 1e-14 < rem < dt_step:
            dt_step = rem
        dt_step = max(dt_step, adaptive.dt_min)
        clamped = dt_step < dt * 0.99
        if len(ws.active) == ws.B:
            # no psi copies: psi_old moves to slot 2, both cores start from it and leave
            # their result in slot 0; psi_strang moves to slot 3
            ws.swap_slots(2, 0)
            strang_core(ws, dt_step, src_slot=2)
            ws.swap_slots(3, 0)
            yoshida_core(ws, dt_step, src_slot=2)
        else:                              # (inactive instances keep their slots in place)
            ws.copy_slot(2, 0)             # psi_old
            strang_core(ws, dt_step)
            ws.copy_slot(3, 0)             # psi_strang
            ws.copy_slot(0, 2)
            yoshida_core(ws, dt_step)
        d = ws.diff(0, 3, 0)[0]
        err = float(d[0] / max(d[1], 1e-300))
        if (not clamped) and err > adaptive.tol and dt_step > adaptive.dt_min * 1.01:
            ws.copy_slot(0, 2)
            dt = max(dt_step * max(0.3, 0.9 * (adaptive.tol / err) ** 0.2), adaptive.dt_min)
            n_rej += 1
            continue
        if ws.has_loss:                    # yoshida.jl:64-66
            L.check(L.lib().bec_loss_step(ws.handle, dt_step), "bec_loss_step")
        ws.t[0] += dt_step
        ws.step[0] += 1
        n_acc += 1
        if not clamped:
            fac = min(3.0, 0.9 * (adaptive.tol / err) ** 0.2) if err > 1e-300 else 3.0
            dt = min(max(dt * fac, adaptive.dt_min), adaptive.dt_max)
        if ws.t[0] >= next_save - 1e-14:
            _record_snapshot(ws, res, keep_snapshots)
            if callback is not None:
                callback(ws, n_acc)
            next_save += save_interval
    if not res.times or abs(res.times[-1] - ws.t[0]) > 1e-12:
        _record_snapshot(ws, res, keep_snapshots)
    return dict(result=res, n_accepted=n_acc, n_rejected=n_rej, final_dt=dt)


def run_simulation_adaptive(ws: Workspace, *, adaptive=AdaptiveDtParams(), t_end: float,
                            save_interval: float, callback=None, keep_snapshots=True):
    """run_simulation_adaptive! (adaptive.jl:57-67), real time.  error_mode "step_change"
    (adaptive.jl:69-194): V(dt/2) K(dt) with the closing half potential deferred and merged
    into the next step (FSAL; flushed at save points and at the end), error = L2 change of
    the step, factors 0.9 sqrt(tol/err) in [0.5, 2].  error_mode "richardson"
    (adaptive.jl:203-296): one Strang step of dt against two of dt/2, factors
    0.9 cbrt(tol/err).  The saved states stay on the device (slots 2 and 3)."""
    lib = L.lib()

    def hp(dth):
        ws.sync_zeeman()
        L.check(lib.bec_half_potential(ws.handle, float(dth), 0), "bec_half_potential")

    def l2(a, b):  # sum |a - b|^2 / sum |b|^2 (_wavefunction_l2_change, adaptive.jl:42-50)
        d = ws.diff(a, b, 0)[0]
        return float(d[0] / max(d[1], 1e-300))

    tol = adaptive.tol
    rich = adaptive.error_mode == "richardson"
    dt = min(max(adaptive.dt_init, adaptive.dt_min), adaptive.dt_max)
    res = SimulationResult()
    _record_snapshot(ws, res, keep_snapshots)
    next_save = ws.t[0] + save_interval
    n_acc = n_rej = 0
    fsal, fsal_dt = False, 0.0
    while ws.t[0] < t_end - 1e-14:
        t = ws.t[0]
        dt_step = min(dt, t_end - t)
        rem = next_save - t
        if 1e-14 < rem < dt_step:
            dt_step = rem
        dt_step = max(dt_step, adaptive.dt_min)
        clamped = dt_step < dt * 0.99
        may_reject = (not clamped) and dt_step > adaptive.dt_min * 1.01
        if rich:
            ws.copy_slot(2, 0)                       # psi_saved
            strang_core(ws, dt_step)
            ws.copy_slot(3, 0)                       # psi_full
            ws.copy_slot(0, 2)
            strang_core(ws, dt_step / 2)
            strang_core(ws, dt_step / 2)
            err = l2(3, 0)
            if may_reject and err > tol:
                ws.copy_slot(0, 2)
                dt = max(dt_step * max(0.5, 0.9 * (tol / err) ** (1 / 3)), adaptive.dt_min)
                n_rej += 1
                continue
        else:
            if not clamped:
                ws.copy_slot(2, 0)                   # psi_old
            prev = (fsal, fsal_dt)
            if fsal and abs(fsal_dt - dt_step) < 1e-14:
                hp(dt_step)
            elif fsal:
                hp(fsal_dt / 2)
                hp(dt_step / 2)
            else:
                hp(dt_step / 2)
            fsal = False
            L.check(lib.bec_kinetic(ws.handle, float(dt_step), 0), "bec_kinetic")
            err = 0.0
            if may_reject:
                err = l2(0, 2)
                if err > tol:
                    ws.copy_slot(0, 2)
                    fsal, fsal_dt = prev
                    dt = max(dt_step * max(0.5, 0.9 * math.sqrt(tol / err)), adaptive.dt_min)
                    n_rej += 1
                    continue
            fsal, fsal_dt = True, dt_step
        if ws.has_loss:
            L.check(lib.bec_loss_step(ws.handle, dt_step), "bec_loss_step")
        ws.t[0] += dt_step
        ws.step[0] += 1
        n_acc += 1
        if not clamped:
            if rich:
                fac = min(2.0, 0.9 * (tol / err) ** (1 / 3)) if err > 1e-300 else 2.0
            else:
                if not may_reject:
                    err = l2(0, 2)
                fac = min(2.0, 0.9 * math.sqrt(tol / err)) if err > 1e-300 else 2.0
            dt = min(max(dt * fac, adaptive.dt_min), adaptive.dt_max)
        if ws.t[0] >= next_save - 1e-14:
            if fsal:
                hp(fsal_dt / 2)
                fsal = False
            _record_snapshot(ws, res, keep_snapshots)
            if callback is not None:
                callback(ws, n_acc)
            next_save += save_interval
    if fsal:
        hp(fsal_dt / 2)
    if not res.times or abs(res.times[-1] - ws.t[0]) > 1e-12:
        _record_snapshot(ws, res, keep_snapshots)
    return dict(result=res, n_accepted=n_acc, n_rejected=n_rej, final_dt=dt)


# ---------------------------------------------------------------------------
# Ground state drivers (ground_state.jl)
# ---------------------------------------------------------------------------
def _validate_itp_zeeman(zeeman: ZeemanParams, F, dt):
    """ground_state.jl:31-42."""
    mx = max(abs(v) for v in zeeman_energies(zeeman, F))
    if mx * dt / 4 > ITP_EXPONENT_LIMIT:
        raise ValueError(f"Zeeman p={zeeman.p} with F={F} and dt={dt} causes overflow in "
                         f"imaginary time (exponent={mx * dt / 4:.1f} > {ITP_EXPONENT_LIMIT}).")


def _validate_itp_interactions(interactions, F, dt, psi=None, c_dd=0.0, grid=None):
    """ground_state.jl:46-78."""
    n_peak = float(np.max(np.sum(np.abs(psi) ** 2, axis=0))) if psi is not None else 1.0
    mc = max(abs(interactions.c0), abs(interactions.c1))
    if mc > 1e-30 and mc * n_peak * dt / 4 > ITP_EXPONENT_LIMIT:
        raise ValueError(f"Spin interaction c0={interactions.c0}, c1={interactions.c1} may cause "
                         f"overflow in imaginary time (estimated exponent="
                         f"{mc * n_peak * dt / 4:.1f} > {ITP_EXPONENT_LIMIT}).")
    if c_dd > 0 and c_dd * F * n_peak * dt / 2 > ITP_EXPONENT_LIMIT:
        raise ValueError(f"DDI c_dd={c_dd} with F={F} may cause overflow in imaginary time.")


def _check_itp_overflow(ws: Workspace, step: int, insts=None):
    """ground_state.jl:4-29 on the device-reduced guard values."""
    g = ws.itp_guard()
    for i in (ws.active if insts is None else insts):
        if not np.isfinite(g[i, 0]):
            raise ValueError(f"NaN detected in ITP at step {step}. Likely DDI potential overflow. "
                             "Reduce dt.")
        if ws.enable_ddi:
            exponent = g[i, 1] * ws.params[i].dt / 2
            if exponent > ITP_EXPONENT_LIMIT:
                raise ValueError(f"DDI potential overflow in ITP at step {step}: "
                                 f"max|Phi|={g[i, 1]:.3g}, exponent={exponent:.1f} > "
                                 f"{ITP_EXPONENT_LIMIT}. Reduce dt (current={ws.params[i].dt}).")


def _check_chunk_overflow(ws: Workspace, ch, step: int, insts):
    """ground_state.jl:4-29 on the guard values that bec_itp_chunk returns."""
    for i in insts:
        if ch[i, L.CHUNK_STATUS] != 0.0:
            raise ValueError(f"NaN detected in ITP at step {step}. Likely DDI potential overflow. "
                             "Reduce dt.")
        if ws.enable_ddi:
            exponent = ch[i, L.CHUNK_PHIMAX] * ws.params[i].dt / 2
            if exponent > ITP_EXPONENT_LIMIT:
                raise ValueError(f"DDI potential overflow in ITP at step {step}: "
                                 f"max|Phi|={ch[i, L.CHUNK_PHIMAX]:.3g}, exponent={exponent:.1f} > "
                                 f"{ITP_EXPONENT_LIMIT}. Reduce dt (current={ws.params[i].dt}).")


def _itp_loop(ws: Workspace, n_steps: int, tol: float, save_every: int, target_mz=None):
    """The ITP convergence loop of find_ground_state for every active instance of ws
    (ground_state.jl:132-162).  Instances retire independently when converged.  With
    target_mz the workspace has normalize_every = 0 and the constrained normalisation runs
    first and after every step (ground_state.jl:124-127,143-146)."""
    insts = ws.active
    if target_mz is not None:
        ws.normalize_constrained(target_mz)
    E_prev = {i: float(ws.energy_terms()[i, L.E_TOTAL]) for i in insts}
    ws.copy_slot(1, 0)
    out = {i: dict(converged=False, dE=float("nan"), dpsi=float("nan"), steps=0) for i in insts}
    running = list(insts)
    step = 0
    while step < n_steps and running:
        if step < 10 or target_mz is not None:
            k = 1
        else:
            k = min(save_every - step % save_every, n_steps - step)
        check = (step + k) % save_every == 0
        if check and target_mz is None:
            # steps, energy, dpsi and psi_prev <- psi with one synchronisation
            ch = ws.itp_chunk(k)
            step += k
            if step <= 10:
                _check_chunk_overflow(ws, ch, step, running)
        else:
            split_step(ws, k)
            step += k
            if step <= 10:
                _check_itp_overflow(ws, step, running)
            if target_mz is not None:
                ws.normalize_constrained(target_mz)
            ch = None
            if check:
                et = ws.energy_terms()
                dd = ws.diff(0, 1, 1)
        if check:
            still = []
            for i in running:
                if ch is not None:
                    E = float(ch[i, L.CHUNK_E])
                    dpsi = float(ch[i, L.CHUNK_DPSI])
                else:
                    E = float(et[i, L.E_TOTAL])
                    pmax = dd[i, 1]
                    dpsi = float(dd[i, 0] / pmax) if pmax > 0 else 0.0
                dE = abs(E - E_prev[i])
                out[i].update(dE=dE, dpsi=dpsi, steps=step)
                if dE < tol and dpsi < tol:
                    out[i]["converged"] = True
                else:
                    E_prev[i] = E
                    still.append(i)
            if still != running:
                running = still
                if running:
                    ws.set_active(running)
            if running and ch is None:
                ws.copy_slot(1, 0)
        else:
            for i in running:
                out[i]["steps"] = step
    ws.set_active(insts)
    et = ws.energy_terms()
    for i in insts:
        out[i]["energy"] = float(et[i, L.E_TOTAL])
    return out


def find_ground_state(*, grid: Grid, atom, interactions: InteractionParams,
                      zeeman=ZeemanParams(), potential=NoPotential(), dt=0.001, n_steps=10000,
                      tol=1e-10, initial_state="polar", psi_init=None, enable_ddi=False,
                      c_dd=float("nan"), secular_ddi=False, adaptive_dt=False, dt_max=None,
                      fft_flags=None, target_magnetization=None, precision="fp64"):
    """ground_state.jl:80-163."""
    F = atom.F
    psi0 = init_psi(grid, SpinSystem(F), initial_state) if psi_init is None else np.array(psi_init)
    _validate_itp_zeeman(zeeman, F, dt)
    eff_cdd = c_dd if (enable_ddi and not math.isnan(c_dd)) else 0.0
    _validate_itp_interactions(interactions, F, dt, psi=psi0, c_dd=eff_cdd)
    if adaptive_dt:
        return _find_ground_state_adaptive(
            grid=grid, atom=atom, interactions=interactions, zeeman=zeeman, potential=potential,
            dt=dt, n_steps=n_steps, tol=tol, psi0=psi0, enable_ddi=enable_ddi, c_dd=c_dd,
            secular_ddi=secular_ddi, dt_max=10.0 * dt if dt_max is None else dt_max,
            precision=precision)
    constrained = target_magnetization is not None
    sp = SimParams(dt=dt, n_steps=n_steps, imaginary_time=True,
                   normalize_every=0 if constrained else 1, save_every=max(1, n_steps // 10))
    ws = make_workspace(grid=grid, atom=atom, interactions=interactions, zeeman=zeeman,
                        potential=potential, sim_params=sp, psi_init=psi0, enable_ddi=enable_ddi,
                        c_dd=c_dd, secular_ddi=secular_ddi, precision=precision)
    r = _itp_loop(ws, n_steps, tol, sp.save_every,
                  target_mz=float(target_magnetization) if constrained else None)[0]
    return dict(workspace=ws, converged=r["converged"], energy=r["energy"], dE=r["dE"],
                dpsi=r["dpsi"], steps=r["steps"])


def _find_ground_state_adaptive(*, grid, atom, interactions, zeeman, potential, dt, n_steps, tol,
                                psi0, enable_ddi, c_dd, secular_ddi, dt_max, precision):
    """ground_state.jl:172-228: every max(1, n_steps // 100) steps compare the energy; a
    rise (or NaN) restores psi and halves dt (floor 1e-8), otherwise test convergence and
    grow dt by 10 % up to dt_max.  psi backup stays on the device (slot 1)."""
    check_every = max(1, n_steps // 100)
    sp = SimParams(dt=dt, n_steps=check_every, imaginary_time=True, normalize_every=1,
                   save_every=check_every)
    ws = make_workspace(grid=grid, atom=atom, interactions=interactions, zeeman=zeeman,
                        potential=potential, sim_params=sp, psi_init=psi0, enable_ddi=enable_ddi,
                        c_dd=c_dd, secular_ddi=secular_ddi, precision=precision)
    cur = float(dt)
    E_prev = total_energy(ws)
    converged = False
    total = 0
    dE = dpsi = float("nan")
    while total < n_steps:
        ws.copy_slot(1, 0)
        for i in range(check_every):
            split_step(ws, 1)
            if i == 0 and total == 0:
                _check_itp_overflow(ws, 1)
        total += check_every
        E = total_energy(ws)
        if math.isnan(E) or E > E_prev:
            ws.copy_slot(0, 1)
            cur = max(cur * 0.5, 1e-8)
            ws.set_dt(0, cur)
        else:
            dE = abs(E - E_prev)
            d = ws.diff(0, 1, 1)[0]
            dpsi = float(d[0] / d[1]) if d[1] > 0 else 0.0
            if dE < tol and dpsi < tol:
                converged = True
                break
            E_prev = E
            new_dt = min(cur * 1.1, dt_max)
            if new_dt != cur:
                cur = new_dt
                ws.set_dt(0, cur)
    return dict(workspace=ws, converged=converged, energy=total_energy(ws), dE=dE, dpsi=dpsi,
                steps=total, final_dt=cur)


def find_ground_state_multistart(*, initial_states=("polar", "ferromagnetic", "uniform",
                                                    "antiferromagnetic"),
                                 n_random=3, seed=42, grid, atom, interactions, **kw):
    """ground_state.jl:239-271: the solves run one after another, each on its own
    workspace; best = first minimum of the energies (argmin)."""
    results = []
    for state in initial_states:
        if state == "random":
            for i in range(1, n_random + 1):
                psi0 = init_psi(grid, SpinSystem(atom.F), "random", seed=seed + i)
                r = find_ground_state(grid=grid, atom=atom, interactions=interactions,
                                      psi_init=psi0, **kw)
                results.append(dict(initial_state="random", idx=i, **r))
        else:
            r = find_ground_state(grid=grid, atom=atom, interactions=interactions,
                                  initial_state=state, **kw)
            results.append(dict(initial_state=state, idx=0, **r))
    best = min(results, key=lambda r: r["energy"])   # first minimum, like argmin
    return dict(workspace=best["workspace"], converged=best["converged"], energy=best["energy"],
                initial_state=best["initial_state"], all_results=results)


def scan_continuation(*, param_values, make_interactions, grid, atom, initial_state="polar",
                      energy_jump_threshold=0.1, n_steps_continuation=500, n_steps_fresh=5000,
                      **kw):
    """ground_state.jl:328-378.  Each record is the reference's (param, energy, converged,
    phase, psi) plus `fallback`: whether the multistart search replaced the continued
    solve after an energy jump."""
    results = []
    prev_psi = None
    prev_E = float("nan")
    for val in param_values:
        inter = make_interactions(val)
        if prev_psi is not None:
            r = find_ground_state(grid=grid, atom=atom, interactions=inter, psi_init=prev_psi.copy(),
                                  n_steps=n_steps_continuation, **kw)
        else:
            r = find_ground_state(grid=grid, atom=atom, interactions=inter,
                                  initial_state=initial_state, n_steps=n_steps_fresh, **kw)
        fallback = False
        if (not math.isnan(prev_E)
                and abs(r["energy"] - prev_E) / max(abs(prev_E), 1e-30) > energy_jump_threshold):
            r["workspace"].close()
            r = find_ground_state_multistart(grid=grid, atom=atom, interactions=inter,
                                             n_steps=n_steps_fresh, **kw)
            fallback = True
        psi = r["workspace"].get_psi(0)
        ph = classify_phase(psi, atom.F, grid)
        results.append(dict(param=val, energy=r["energy"], converged=r["converged"],
                            phase=ph["phase"], psi=psi.copy(), fallback=fallback))
        prev_psi = psi
        prev_E = r["energy"]
    return results
