"""Batched ground-state solves: many independent find_ground_state problems of mixed
shapes, grouped into one workspace per (grid, spin, precision, dt, n_steps, tol) with one
instance per solve, advanced together on the device.

The per-solve semantics are exactly those of find_ground_state (ground_state.jl:80-163):
ITP split steps with normalisation every step, the overflow guard on the first 10 steps, a
convergence test every max(1, n_steps // 10) steps on |dE| and max|dpsi|/max|psi|, and the
energy recomputed at exit.  Each instance retires when it converges; the kernels only visit
the active instances.  Workspaces of different shapes run on their own HIP streams; the
host issues their steps back to back and only waits on a workspace at its convergence
checks, so the groups overlap on the device.

find_ground_state_multistart (ground_state.jl:239-271) over a list of initial states is
the special case of one group; `multistart_batched` returns the same record as the
sequential driver.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field, replace

import numpy as np

from . import _lib as L
from .api import Workspace, _check_itp_overflow, _validate_itp_interactions, _validate_itp_zeeman
from .physics import (Grid, HarmonicTrap, InteractionParams, NoPotential, SimParams, SpinSystem,
                      ZeemanParams, init_psi)


class WorkspaceGroup:
    """bec_group_*: several workspaces of different grid shapes (same precision, spin,
    dimension and time kind) advanced together, each position of their common split-step
    launch sequence issued as one kernel launch over all of them.  The group keeps
    references to its members and must be closed before they are."""

    def __init__(self, workspaces):
        self.members = list(workspaces)
        arr = (C.c_void_p * len(self.members))(*[w.handle for w in self.members])
        h = C.c_void_p()
        L.check(L.lib().bec_group_create(arr, len(self.members), C.byref(h)), "bec_group_create")
        self._h = h

    def split_step(self, n_steps, graph=True):
        L.check(L.lib().bec_group_split_step(self._h, int(n_steps), int(bool(graph))),
                "bec_group_split_step")

    @property
    def grouped(self):
        """Whether the last split_step issued grouped launches."""
        return bool(L.lib().bec_group_grouped(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None:
            L.lib().bec_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ITPJob:
    """One find_ground_state problem."""
    grid: Grid
    F: int
    interactions: InteractionParams
    psi0: np.ndarray
    zeeman: ZeemanParams = field(default_factory=ZeemanParams)
    potential: object = field(default_factory=NoPotential)
    dt: float = 0.001
    n_steps: int = 10000
    tol: float = 1e-10
    tag: object = None


def _group_key(job: ITPJob, precision):
    return (tuple(job.grid.config.n_points), tuple(job.grid.config.box_size), job.F, precision,
            float(job.dt), int(job.n_steps), float(job.tol))


class _Group:
    """The instances of one workspace and their ITP loop state."""

    def __init__(self, jobs, idx, precision, graphs=False, autotune=True):
        j0 = jobs[0]
        self.graphs = graphs
        self.idx = idx
        self.jobs = jobs
        self.save_every = max(1, j0.n_steps // 10)
        self.n_steps = j0.n_steps
        self.tol = j0.tol
        sp = SimParams(dt=j0.dt, n_steps=j0.n_steps, imaginary_time=True, normalize_every=1,
                       save_every=self.save_every)
        self.ws = Workspace(grid=j0.grid, F=j0.F, n_instances=len(jobs), sim_params=sp,
                            precision=precision, autotune=autotune)
        for i, j in enumerate(jobs):
            self.ws.set_instance(i, interactions=j.interactions, zeeman=j.zeeman,
                                 potential=j.potential, dt=j.dt)
            self.ws.set_psi(np.asarray(j.psi0, dtype=complex), i)
        self.step = 0
        self.running = list(range(len(jobs)))
        et = self.ws.energy_terms()
        self.E_prev = {i: float(et[i, L.E_TOTAL]) for i in self.running}
        self.ws.copy_slot(1, 0)
        self.out = {i: dict(converged=False, dE=float("nan"), dpsi=float("nan"), steps=0)
                    for i in self.running}
        self.t_done = {}

    @property
    def done(self):
        return not self.running or self.step >= self.n_steps

    def next_k(self):
        """Steps up to the next check point (1 at a time for the first 10)."""
        if self.step < 10:
            return 1
        return min(self.save_every - self.step % self.save_every, self.n_steps - self.step)

    def issue(self):
        """Queue the steps up to the next check point."""
        k = self.next_k()
        if self.graphs:
            L.check(L.lib().bec_split_step_graph(self.ws.handle, int(k)), "bec_split_step_graph")
        else:
            L.check(L.lib().bec_split_step(self.ws.handle, int(k)), "bec_split_step")
        self.step += k
        return k

    def check(self, t0):
        """Overflow guard / convergence test after issue(); waits on this workspace only."""
        if self.step <= 10:
            _check_itp_overflow(self.ws, self.step, self.running)
        if self.step % self.save_every == 0:
            et = self.ws.energy_terms()
            dd = self.ws.diff(0, 1, 1)
            still = []
            for i in self.running:
                E = float(et[i, L.E_TOTAL])
                dE = abs(E - self.E_prev[i])
                pmax = dd[i, 1]
                dpsi = float(dd[i, 0] / pmax) if pmax > 0 else 0.0
                self.out[i].update(dE=dE, dpsi=dpsi, steps=self.step)
                if dE < self.tol and dpsi < self.tol:
                    self.out[i]["converged"] = True
                    self.t_done[i] = time.perf_counter() - t0
                else:
                    self.E_prev[i] = E
                    still.append(i)
            if still != self.running:
                self.running = still
                if still:
                    self.ws.set_active(still)
            if self.running:
                self.ws.copy_slot(1, 0)
        else:
            for i in self.running:
                self.out[i]["steps"] = self.step
        if self.done:
            for i in self.running:
                self.t_done.setdefault(i, time.perf_counter() - t0)

    def finish(self, keep_psi):
        self.ws.set_active(range(len(self.jobs)))
        et = self.ws.energy_terms()
        res = []
        for i in range(len(self.jobs)):
            r = dict(self.out[i])
            r["energy"] = float(et[i, L.E_TOTAL])
            r["time_s"] = self.t_done.get(i, float("nan"))
            if keep_psi:
                r["psi"] = self.ws.get_psi(i)
            res.append(r)
        return res


def solve_batch(jobs, *, precision="fp64", keep_psi=False, graphs=False, grouped=False,
                autotune=None):
    """Run every job to convergence or n_steps.  Returns (results in job order, stats).
    graphs: issue each group's steps between check points as captured HIP graphs.
    grouped: when the live groups are due at the same check point, advance all of them
    with grouped launches (WorkspaceGroup: one launch per pass over every shape; captured
    as one graph with `graphs`).  A grouped x-pass launch needs one block size for every
    shape, which the default tiles give (the same number of sites per block), so grouped
    runs keep them: autotune (per-shape x tiles, default on) is off with `grouped`.
    Results are bitwise the same in every mode with the same tiles."""
    if autotune is None:
        autotune = not grouped
    for j in jobs:
        _validate_itp_zeeman(j.zeeman, j.F, j.dt)
        _validate_itp_interactions(j.interactions, j.F, j.dt, psi=j.psi0)
    groups = {}
    for k, j in enumerate(jobs):
        groups.setdefault(_group_key(j, precision), []).append(k)
    t0 = time.perf_counter()
    gs = [_Group([jobs[k] for k in idx], idx, precision, graphs, autotune)
          for idx in groups.values()]
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    live = list(gs)
    wgroup, wmembers = None, None
    n_grouped = 0
    while live:
        ks = {g.next_k() for g in live}
        if grouped and len(live) > 1 and len(ks) == 1:
            if wmembers != live:  # the group follows the live set
                if wgroup is not None:
                    wgroup.close()
                wgroup, wmembers = WorkspaceGroup([g.ws for g in live]), list(live)
            k = ks.pop()
            wgroup.split_step(k, graph=graphs)
            n_grouped += int(wgroup.grouped)
            for g in live:
                g.step += k
        else:
            for g in live:      # queue work on every stream first ...
                g.issue()
        for g in live:          # ... then wait on each group at its check point
            g.check(t0)
        live = [g for g in live if not g.done]
    if wgroup is not None:
        wgroup.close()
    for g in gs:
        g.ws.synchronize()
    wall = time.perf_counter() - t0
    results = [None] * len(jobs)
    for g in gs:
        for i, r in enumerate(g.finish(keep_psi)):
            results[g.idx[i]] = r
    for g in gs:
        g.ws.close()
    stats = dict(wall_s=wall, setup_s=t_setup, n_groups=len(gs), solves=len(jobs),
                 solves_per_s=len(jobs) / wall,
                 group_sizes=[len(g.jobs) for g in gs],
                 group_steps=[g.step for g in gs], grouped_intervals=n_grouped)
    return results, stats


def solve_batch_refined(jobs, *, refine_steps=2000, keep_psi=False, graphs=False, grouped=False,
                        autotune=None):
    """FP32-bulk solves finished in FP64 (iterative refinement of the FP32 fixed point).

    Every job runs in FP32-bulk to convergence or n_steps (solve_batch); its state, widened
    to FP64, then runs exactly refine_steps more ITP steps in FP64 (same dt, the overflow
    guard and energies of find_ground_state, no early stop).  An FP32 iteration stalls once
    a mode's per-step change falls below half an ulp of psi (the slowest mode of an untrapped
    box freezes near 1e-4 relative amplitude, 1.45e-7 in the energy); the FP64 steps let it
    decay.  Returns (FP64 results in job order, stats); each result also carries the FP32
    phase's steps, convergence and energy."""
    r32, s32 = solve_batch(jobs, precision="fp32", keep_psi=True, graphs=graphs,
                           grouped=grouped, autotune=autotune)
    jobs64 = [replace(j, psi0=np.asarray(r["psi"], dtype=np.complex128), n_steps=refine_steps,
                      tol=0.0) for j, r in zip(jobs, r32)]
    r64, s64 = solve_batch(jobs64, precision="fp64", keep_psi=keep_psi, graphs=graphs,
                           grouped=grouped, autotune=autotune)
    out = []
    for a, b in zip(r32, r64):
        r = dict(b)
        r["steps_fp32"], r["converged_fp32"], r["energy_fp32"] = a["steps"], a["converged"], a["energy"]
        r["converged"] = a["converged"]
        r["time_s"] = a["time_s"] + s64["wall_s"]
        out.append(r)
    wall = s32["wall_s"] + s64["wall_s"]
    stats = dict(wall_s=wall, solves=len(jobs), solves_per_s=len(jobs) / wall,
                 refine_steps=refine_steps, fp32=s32, fp64_refine=s64)
    return out, stats


def multistart_batched(*, initial_states=("polar", "ferromagnetic", "uniform",
                                          "antiferromagnetic"),
                       n_random=3, seed=42, grid, atom, interactions, zeeman=ZeemanParams(),
                       potential=NoPotential(), dt=0.001, n_steps=10000, tol=1e-10,
                       precision="fp64"):
    """find_ground_state_multistart (ground_state.jl:239-271) with all starts in one batch;
    best = first minimum of the energies, as the sequential driver."""
    starts = []
    for state in initial_states:
        if state == "random":
            for i in range(1, n_random + 1):
                starts.append(("random", i, init_psi(grid, SpinSystem(atom.F), "random",
                                                     seed=seed + i)))
        else:
            starts.append((state, 0, init_psi(grid, SpinSystem(atom.F), state)))
    jobs = [ITPJob(grid=grid, F=atom.F, interactions=interactions, psi0=p, zeeman=zeeman,
                   potential=potential, dt=dt, n_steps=n_steps, tol=tol, tag=(s, i))
            for s, i, p in starts]
    res, stats = solve_batch(jobs, precision=precision, keep_psi=True)
    best = min(range(len(res)), key=lambda k: res[k]["energy"])
    all_results = [dict(initial_state=jobs[k].tag[0], idx=jobs[k].tag[1], **res[k])
                   for k in range(len(res))]
    return dict(converged=res[best]["converged"], energy=res[best]["energy"],
                initial_state=jobs[best].tag[0], psi=res[best]["psi"],
                all_results=all_results, stats=stats)


def scan_continuation_batched(chains, *, grid, atom, initial_state="polar",
                              energy_jump_threshold=0.1, n_steps_continuation=500,
                              n_steps_fresh=5000, zeeman=ZeemanParams(), potential=NoPotential(),
                              dt=0.001, tol=1e-10, precision="fp64", keep_psi=True):
    """scan_continuation (ground_state.jl:328-378) for several independent chains at once.

    chains: a list of (param_values, make_interactions).  Point i of every chain that has
    one is solved in the same batch (solve_batch), continued from that chain's previous
    ground state (n_steps_continuation) or, for the first point, from `initial_state`
    (n_steps_fresh); the chains whose energy jumped by more than energy_jump_threshold
    (relative) are solved again by multistart over the reference's four named states, all
    such fallbacks of the point in one more batch.  Each chain's records are those of
    scan_continuation: param, energy, converged, phase, psi, fallback."""
    from .physics import classify_phase
    sys_ = SpinSystem(atom.F)
    named = ("polar", "ferromagnetic", "uniform", "antiferromagnetic")
    psi_named = {s: init_psi(grid, sys_, s) for s in named + (initial_state,)}
    out = [[] for _ in chains]
    prev_psi = [None] * len(chains)
    prev_E = [float("nan")] * len(chains)
    n_points = max(len(vals) for vals, _ in chains)
    for i in range(n_points):
        live = [c for c, (vals, _) in enumerate(chains) if i < len(vals)]
        jobs = []
        for c in live:
            vals, mk = chains[c]
            fresh = prev_psi[c] is None
            jobs.append(ITPJob(grid=grid, F=atom.F, interactions=mk(vals[i]),
                               psi0=psi_named[initial_state] if fresh else prev_psi[c].copy(),
                               zeeman=zeeman, potential=potential, dt=dt,
                               n_steps=n_steps_fresh if fresh else n_steps_continuation,
                               tol=tol, tag=c))
        res, _ = solve_batch(jobs, precision=precision, keep_psi=True)
        point = {c: r for c, r in zip(live, res)}
        jumped = [c for c in live
                  if not math.isnan(prev_E[c])
                  and abs(point[c]["energy"] - prev_E[c]) / max(abs(prev_E[c]), 1e-30)
                  > energy_jump_threshold]
        if jumped:
            fjobs = [ITPJob(grid=grid, F=atom.F, interactions=chains[c][1](chains[c][0][i]),
                            psi0=psi_named[s], zeeman=zeeman, potential=potential, dt=dt,
                            n_steps=n_steps_fresh, tol=tol, tag=(c, s))
                     for c in jumped for s in named]
            fres, _ = solve_batch(fjobs, precision=precision, keep_psi=True)
            for c in jumped:
                rs = [r for j, r in zip(fjobs, fres) if j.tag[0] == c]
                best = min(range(len(rs)), key=lambda k: rs[k]["energy"])   # first minimum
                point[c] = dict(rs[best], fallback=True)
        for c in live:
            r = point[c]
            psi = r["psi"]
            out[c].append(dict(param=chains[c][0][i], energy=r["energy"],
                               converged=r["converged"],
                               phase=classify_phase(psi, atom.F, grid)["phase"],
                               psi=psi.copy() if keep_psi else None,
                               fallback=bool(r.get("fallback", False))))
            prev_psi[c] = psi
            prev_E[c] = r["energy"]
    return out
