"""Continuous-batching request server for ground-state (ITP) and real-time solves.

Requests of mixed dimension, spin and grid size arrive over time.  Each one is admitted
into a free instance slot of a device-resident workspace of its shape (a *pool*).  A pool
is created on demand when the HBM budget allows it (its size is bec_workspace_bytes) and
an idle pool is destroyed when memory is needed for another shape.  Every pool advances
all of its active slots together on its own HIP stream; slots at different points of
their solves coexist because a chunk of steps never crosses any slot's next event
(convergence check or end of run), and chunks are sized to about `chunk_ms` of device
time so that requests arriving meanwhile join at the next chunk boundary.  The host
issues the chunks of all busy pools before waiting on any, so pools of different shapes
overlap on the device.

Per request the arithmetic does not depend on the mix: the kernels work on one instance
per grid row and partition every reduction by the grid shape only, an instance's slot
index changes no operand, and split steps give the same results however they are grouped
into launches.  `solve_alone` runs a request on a single-slot pool for comparison.

Request semantics:
  * "itp": find_ground_state (ground_state.jl:80-163): imaginary-time split steps with
    normalisation after every step, a convergence test every max(1, n_steps // 10) steps
    on |dE| < tol and max|dpsi| / max|psi| < tol, the energy of the final state.  The
    overflow guard of the first ten steps (ground_state.jl:4-29) becomes a check for a
    non-finite energy at every check point; such a request ends with `error` set.
  * "rt": n_steps real-time split_step! (split_step.jl:10-35) and the final energy.
Both start from init_psi(state) (initialization.jl:1-67) in a harmonic trap.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .api import Workspace, _validate_itp_interactions
from .physics import (GridConfig, HarmonicTrap, InteractionParams, SimParams, SpinSystem,
                      SplitMix64, ZeemanParams, init_psi, make_grid)


@dataclass
class Request:
    rid: int
    kind: str                 # "itp" or "rt"
    n: tuple                  # grid points per axis (any even side)
    L: tuple                  # box length per axis
    F: int
    c0: float
    c1: float
    n_steps: int
    dt: float
    tol: float = 1e-8
    omega: float = 1.0
    state: str = "polar"
    seed: int = 0
    arrival: float = 0.0      # seconds after the server start
    precision: str = "fp64"

    @property
    def sites(self):
        return int(np.prod(self.n))


@dataclass
class Result:
    rid: int
    energy: float
    steps: int
    converged: bool
    arrival: float
    admitted: float
    finished: float
    error: str = None
    psi: np.ndarray = None

    @property
    def latency(self):
        return self.finished - self.arrival


def _pool_key(r: Request):
    return (r.kind, tuple(r.n), tuple(r.L), r.F, r.precision)


def pool_bytes(grid, F, itp, precision, slots):
    """Device bytes of a pool: bec_workspace_bytes plus, for ITP, the saved state of the
    convergence test (one more psi-sized buffer)."""
    b = Workspace.bytes_for(grid, F, slots, precision)
    if itp:
        b += slots * (2 * F + 1) * int(np.prod(grid.shape)) * (8 if precision == "fp32" else 16)
    return b


class _Pool:
    """A workspace of one shape with `slots` instances, each free or holding a request."""

    def __init__(self, key, grid, F, itp, precision, slots):
        sp = SimParams(dt=1e-3, n_steps=1, imaginary_time=itp, normalize_every=1 if itp else 0)
        self.key = key
        self.itp = itp
        self.grid = grid
        self.F = F
        self.ws = Workspace(grid=grid, F=F, n_instances=slots, sim_params=sp, precision=precision)
        self.bytes = pool_bytes(grid, F, itp, precision, slots)
        self.free = list(range(slots))
        self.slot = {}            # slot -> per-request state
        self.ms_per_step = None
        self.k = 0
        self.t_issue = 0.0
        self.steps_issued = 0
        self.in_flight = False

    @property
    def active(self):
        return sorted(self.slot)


class Server:
    """Single host thread; device-resident pools; continuous batching.  graphs: issue each
    chunk as a captured HIP graph (bec_split_step_graph), chunk lengths rounded down to
    powers of two so that each pool replays a few cached graphs."""

    def __init__(self, *, hbm_budget_bytes=200 << 30, slots_per_pool=8, chunk_ms=20.0,
                 max_chunk=512, keep_psi=False, graphs=False, max_in_flight=8):
        self.max_in_flight = int(max_in_flight)
        self.budget = int(hbm_budget_bytes)
        self.slots_per_pool = int(slots_per_pool)
        self.chunk_ms = float(chunk_ms)
        self.max_chunk = int(max_chunk)
        self.keep_psi = keep_psi
        self.graphs = graphs
        self.pools = {}
        self.used = 0
        self.peak_used = 0
        self.grids = {}
        self.pools_created = 0
        self.host_issue_s = 0.0

    # -- admission -------------------------------------------------------------
    def _grid(self, r):
        k = (tuple(r.n), tuple(r.L))
        if k not in self.grids:
            self.grids[k] = make_grid(GridConfig(tuple(r.n), tuple(float(v) for v in r.L)))
        return self.grids[k]

    def pool_bytes(self, r):
        return pool_bytes(self._grid(r), r.F, r.kind == "itp", r.precision, self.slots_per_pool)

    def _pool_for(self, r):
        """The pool that can take `r` now, creating it within the HBM budget; else None."""
        key = _pool_key(r)
        p = self.pools.get(key)
        if p is not None:
            return p if (p.free and not p.in_flight) else None
        need = self.pool_bytes(r)
        if need > self.budget:
            raise ValueError(f"request {r.rid}: a pool of {need} bytes exceeds the HBM budget")
        if self.used + need > self.budget:
            for k, q in list(self.pools.items()):   # reclaim idle pools, oldest first
                if not q.slot and not q.in_flight:
                    self.used -= q.bytes
                    q.ws.close()
                    del self.pools[k]
                    if self.used + need <= self.budget:
                        break
        if self.used + need > self.budget:
            return None
        p = _Pool(key, self._grid(r), r.F, r.kind == "itp", r.precision, self.slots_per_pool)
        self.pools[key] = p
        self.used += p.bytes
        self.peak_used = max(self.peak_used, self.used)
        self.pools_created += 1
        return p

    def _admit(self, p: _Pool, r: Request, now, results):
        psi = init_psi(p.grid, SpinSystem(r.F), r.state, seed=r.seed)
        inter = InteractionParams(r.c0, r.c1)
        if r.kind == "itp":
            try:
                _validate_itp_interactions(inter, r.F, r.dt, psi=psi)
            except ValueError as e:
                results.append(Result(rid=r.rid, energy=float("nan"), steps=0, converged=False,
                                      arrival=r.arrival, admitted=now, finished=now, error=str(e)))
                return
        s = p.free.pop(0)
        p.ws.set_instance(s, interactions=inter, zeeman=ZeemanParams(),
                          potential=HarmonicTrap(*([r.omega] * len(r.n))), dt=r.dt)
        p.ws.set_loss(s, None)
        p.ws.set_psi(psi, s)
        st = dict(req=r, step=0, admitted=now, converged=False)
        if r.kind == "itp":
            st["save_every"] = max(1, r.n_steps // 10)
            st["next"] = min(st["save_every"], r.n_steps)
            p.ws.set_active([s])
            st["E_prev"] = float(p.ws.energy_terms()[s, L.E_TOTAL])
            p.ws.copy_slot(1, 0)
        else:
            st["next"] = r.n_steps
        p.slot[s] = st
        p.ws.set_active(p.active)

    # -- stepping ------------------------------------------------------------------
    @staticmethod
    def _remaining_ms(p: _Pool):
        """Work left in the pool for the scheduling order: its longest slot's remaining steps
        times a static per-step cost (a launch overhead plus a term in the sites and
        components; the measured pace is not used, because a pool sharing the device with
        large ones measures their time as its own)."""
        steps = max(st["req"].n_steps - st["step"] for st in p.slot.values())
        return steps * (0.02 + 2e-6 * int(np.prod(p.grid.shape)) * (2 * p.F + 1))

    def _issue(self, p: _Pool):
        """Queue one chunk on the pool's stream: up to the nearest event of its slots, and
        about chunk_ms of the pool's measured pace (a few steps while it is unmeasured)."""
        k = min(st["next"] - st["step"] for st in p.slot.values())
        k = min(k, max(1, int(self.chunk_ms / p.ms_per_step)) if p.ms_per_step else 4)
        k = max(1, min(k, self.max_chunk))
        t = time.perf_counter()
        if self.graphs:
            # chunks of 2^j steps, so that a pool replays a few captured graphs
            k = 1 << (k.bit_length() - 1)
            L.check(L.lib().bec_split_step_graph(p.ws.handle, int(k)), "bec_split_step_graph")
        else:
            L.check(L.lib().bec_split_step(p.ws.handle, int(k)), "bec_split_step")
        self.host_issue_s += time.perf_counter() - t
        for st in p.slot.values():
            st["step"] += k
        p.k = k
        p.steps_issued += k
        p.t_issue = t
        p.in_flight = True

    def _retire(self, p, s, et, t0, results, error=None):
        st = p.slot.pop(s)
        r = st["req"]
        results.append(Result(rid=r.rid, energy=float(et[s, L.E_TOTAL]), steps=st["step"],
                              converged=st["converged"], arrival=r.arrival,
                              admitted=st["admitted"], finished=time.perf_counter() - t0,
                              error=error, psi=p.ws.get_psi(s) if self.keep_psi else None))
        p.free.append(s)

    def _events(self, p: _Pool, t0, results):
        """The pool's chunk has finished: convergence checks and retirements of its due slots."""
        p.in_flight = False
        ms = (time.perf_counter() - p.t_issue) * 1e3 / p.k
        p.ms_per_step = ms if p.ms_per_step is None else 0.7 * p.ms_per_step + 0.3 * ms
        due = [s for s, st in p.slot.items() if st["step"] >= st["next"]]
        if not due:
            return
        p.ws.set_active(due)
        et = p.ws.energy_terms()
        if p.itp:
            dd = p.ws.diff(0, 1, 1)
            keep = []
            for s in due:
                st = p.slot[s]
                r = st["req"]
                E = float(et[s, L.E_TOTAL])
                if not math.isfinite(E):
                    self._retire(p, s, et, t0, results, error="non-finite energy in ITP; reduce dt")
                    continue
                if st["step"] % st["save_every"] == 0:     # ground_state.jl:139-157
                    dE = abs(E - st["E_prev"])
                    dpsi = float(dd[s, 0] / dd[s, 1]) if dd[s, 1] > 0 else 0.0
                    if dE < r.tol and dpsi < r.tol:
                        st["converged"] = True
                        self._retire(p, s, et, t0, results)
                        continue
                    st["E_prev"] = E
                if st["step"] >= r.n_steps:
                    self._retire(p, s, et, t0, results)
                else:
                    keep.append(s)
                    st["next"] = min(st["step"] + st["save_every"], r.n_steps)
            if keep:
                p.ws.set_active(keep)
                p.ws.copy_slot(1, 0)
        else:
            for s in due:
                self._retire(p, s, et, t0, results)
        if p.slot:
            p.ws.set_active(p.active)

    def run(self, requests, realtime=True):
        """Serve `requests`.  realtime: each request becomes visible at its arrival time on
        the wall clock (latency = finish - arrival); otherwise one request is released per
        loop iteration.  The loop polls the pools' streams without blocking: a pool whose
        chunk has finished is checked, takes waiting requests and is issued its next chunk
        while the other pools' chunks still run.  Returns (results in completion order,
        stats)."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        queue, results = [], []
        self.host_issue_s = 0.0
        t0 = time.perf_counter()
        it = 0
        while pending or queue or any(p.slot for p in self.pools.values()):
            progressed = False
            for p in list(self.pools.values()):
                if p.in_flight and p.ws.query():
                    self._events(p, t0, results)
                    progressed = True
            now = time.perf_counter() - t0
            if realtime:
                while pending and pending[0].arrival <= now:
                    queue.append(pending.pop(0))
            elif pending:
                queue.append(pending.pop(0))
            waiting = []
            for r in queue:
                p = self._pool_for(r)
                if p is None:
                    waiting.append(r)
                else:
                    self._admit(p, r, time.perf_counter() - t0, results)
                    progressed = True
            queue = waiting
            # At most max_in_flight pools have a chunk on the device (the process has a few
            # hardware queues; more streams only interleave their chunks), chosen shortest
            # remaining work first, so that small requests are not queued behind the chunks
            # of large ones.  Results do not depend on this order (chunking does not change
            # an instance's arithmetic).
            busy = sum(1 for p in self.pools.values() if p.in_flight)
            ready = [p for p in self.pools.values() if p.slot and not p.in_flight]
            for p in sorted(ready, key=self._remaining_ms):
                if busy >= self.max_in_flight:
                    break
                self._issue(p)
                busy += 1
                progressed = True
            if not any(p.in_flight for p in self.pools.values()):
                if queue:
                    raise RuntimeError("requests wait for memory that no pool will release")
                if pending and realtime:
                    time.sleep(max(0.0, min(0.005, pending[0].arrival - now)))
            elif not progressed:
                time.sleep(5e-5)
            it += 1
        wall = time.perf_counter() - t0
        for p in self.pools.values():
            p.ws.close()
        self.pools, self.used = {}, 0
        return results, dict(wall_s=wall, iterations=it, pools_created=self.pools_created,
                             peak_pool_bytes=self.peak_used, host_issue_s=self.host_issue_s)


def solve_alone(r: Request, keep_psi=True):
    """The same request on a single-slot pool of its own, nothing else on the device."""
    srv = Server(slots_per_pool=1, keep_psi=keep_psi)
    q = Request(**{**r.__dict__, "arrival": 0.0})
    res, _ = srv.run([q], realtime=False)
    return res[0]


# The stream's side ranges: every even side in {64 ... 256} for 1-D and 2-D requests and in
# {48 ... 128} for 3-D (sides without a built radix plan run the runtime-length passes).
SIDES12 = tuple(range(64, 257, 2))
SIDES3 = tuple(range(48, 129, 2))


def poisson_mix(n_req, *, rate=200.0, seed=42, steps=(500, 5000), sides12=SIDES12,
                sides3=SIDES3, precision="fp64"):
    """A seeded request stream: exponential inter-arrival times of mean 1/rate; dimension
    1/2/3 with probability 0.5/0.35/0.15; F = 1/2/6 with probability 0.6/0.25/0.15; side
    uniform over `sides12` (1-D, 2-D) or `sides3` (3-D); c0 ~ U[5, 20], c1 ~ U[-1, 1];
    n_steps ~ U[steps]; ITP or real time with probability 1/2; harmonic trap omega = 1,
    box 20 per axis; polar / ferromagnetic / uniform initial states; ITP dt 0.005 with
    tol 1e-8, real-time dt 0.001."""
    g = SplitMix64(seed)
    out = []
    t = 0.0
    for i in range(n_req):
        u = g.uniform(9)
        t += -math.log(u[0]) / rate
        nd = 1 if u[1] < 0.5 else (2 if u[1] < 0.85 else 3)
        F = 1 if u[2] < 0.6 else (2 if u[2] < 0.85 else 6)
        sides = sides3 if nd == 3 else sides12
        n = sides[min(len(sides) - 1, int(u[3] * len(sides)))]
        kind = "itp" if u[4] < 0.5 else "rt"
        n_steps = int(steps[0] + u[5] * (steps[1] - steps[0]))
        state = ("polar", "ferromagnetic", "uniform")[min(2, int(u[8] * 3))]
        out.append(Request(rid=i, kind=kind, n=(n,) * nd, L=(20.0,) * nd, F=F,
                           c0=5.0 + 15.0 * u[6], c1=-1.0 + 2.0 * u[7], n_steps=n_steps,
                           dt=0.005 if kind == "itp" else 0.001, tol=1e-8, state=state,
                           seed=seed + i, arrival=t, precision=precision))
    return out
