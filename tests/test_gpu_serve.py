"""Continuous-batching server (becgpu.serve): each request's result is the same alone and in
the mix (bitwise: the kernels give an instance the same arithmetic whatever else shares its
workspace, and split steps are chunking-invariant), and a served ITP request equals
find_ground_state on the same inputs."""
import math

import numpy as np
import pytest

import becgpu as G
from becgpu import api as A
from becgpu.serve import Request, Server, solve_alone

pytestmark = pytest.mark.gpu


def _mix(precision):
    """Twelve requests over five pool shapes; with two slots per pool some requests queue
    for a slot and reuse one freed by an earlier request."""
    spec = [  # kind, n, F, c0, c1, n_steps, state
        ("itp", (64,), 1, 10.0, -0.5, 300, "polar"),
        ("rt", (64,), 1, 12.0, 0.5, 120, "ferromagnetic"),
        ("itp", (64,), 1, 8.0, 0.7, 250, "uniform"),
        ("itp", (64,), 1, 15.0, -0.9, 200, "polar"),
        ("itp", (32, 32), 1, 10.0, 0.3, 180, "ferromagnetic"),
        ("rt", (32, 32), 2, 9.0, -0.4, 90, "uniform"),
        ("itp", (64,), 6, 11.0, 0.2, 160, "ferromagnetic"),
        ("rt", (64,), 1, 6.0, -0.8, 70, "uniform"),
        ("itp", (32, 32), 1, 18.0, -0.2, 220, "uniform"),
        ("rt", (32, 32), 2, 14.0, 0.9, 60, "polar"),
        ("itp", (64,), 6, 7.0, -0.6, 140, "uniform"),
        ("itp", (16, 16, 16), 1, 10.0, 0.4, 100, "polar"),
    ]
    return [Request(rid=i, kind=k, n=n, L=(20.0,) * len(n), F=F, c0=c0, c1=c1, n_steps=ns,
                    dt=0.005 if k == "itp" else 0.001, tol=1e-7, state=st, seed=7 + i,
                    arrival=0.0, precision=precision)
            for i, (k, n, F, c0, c1, ns, st) in enumerate(spec)]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_alone_equals_in_mix(precision):
    reqs = _mix(precision)
    srv = Server(slots_per_pool=2, chunk_ms=0.5, max_chunk=37, keep_psi=True)
    res, stats = srv.run(reqs, realtime=False)
    print(precision, stats)
    assert len(res) == len(reqs)
    by = {r.rid: r for r in res}
    for q in reqs:
        m = by[q.rid]
        a = solve_alone(q)
        print(q.rid, q.kind, q.n, q.F, m.steps, m.converged, m.energy, a.energy)
        assert m.error is None and a.error is None
        assert math.isfinite(m.energy)
        assert (m.steps, m.converged) == (a.steps, a.converged)
        # the tolerance is zero: identical arithmetic in both runs
        assert m.energy == a.energy
        assert np.array_equal(m.psi, a.psi)


def test_served_itp_equals_find_ground_state():
    reqs = [q for q in _mix("fp64") if q.kind == "itp" and len(q.n) == 1][:3]
    srv = Server(slots_per_pool=4, chunk_ms=0.5, max_chunk=23, keep_psi=False)
    res, _ = srv.run(reqs, realtime=False)
    by = {r.rid: r for r in res}
    for q in reqs:
        grid = G.make_grid(G.GridConfig(q.n, q.L))
        atom = G.AtomSpecies("test", 1.0, q.F)
        s = A.find_ground_state(grid=grid, atom=atom, interactions=G.InteractionParams(q.c0, q.c1),
                                potential=G.HarmonicTrap(q.omega), dt=q.dt, n_steps=q.n_steps,
                                tol=q.tol, initial_state=q.state)
        s["workspace"].close()
        assert by[q.rid].energy == s["energy"]
        assert (by[q.rid].steps, by[q.rid].converged) == (s["steps"], s["converged"])


def test_realtime_arrivals_and_budget():
    """Arrivals on the wall clock and an HBM budget that holds one pool at a time: requests
    of a second shape wait until the first pool idles and is reclaimed."""
    reqs = _mix("fp64")[:5]
    for k, q in enumerate(reqs):
        q.arrival = 0.01 * k
    srv = Server(slots_per_pool=2, chunk_ms=1.0, keep_psi=True)
    one = srv.pool_bytes(reqs[4])              # the 2-D ITP pool, the largest of the five
    srv.budget = one + 1
    res, stats = srv.run(reqs, realtime=True)
    print(stats)
    assert sorted(r.rid for r in res) == [q.rid for q in reqs]
    assert stats["peak_pool_bytes"] <= one + 1
    assert stats["pools_created"] >= 3
    for r in res:
        assert r.latency >= 0 and r.finished >= r.admitted >= r.arrival - 1e-9
        q = reqs[r.rid]
        a = solve_alone(q)
        assert r.energy == a.energy and np.array_equal(r.psi, a.psi)


def _stream_slice(precision, seed=7):
    """A seeded slice of the serving workload's stream (poisson_mix: Poisson arrivals at
    200/s, dimension 1/2/3 w.p. 0.5/0.35/0.15, F = 1/2/6 w.p. 0.6/0.25/0.15, sides in the
    configuration's ranges, 500-5000 steps): the first 3-D spin-6 request and the first seven
    1-D/2-D requests whose FP64 oracle run stays under 1.5e8 component-site-steps."""
    from becgpu.serve import poisson_mix
    f6, rest = None, []
    for q in poisson_mix(300, seed=seed, precision=precision):
        if len(q.n) == 3 and q.F == 6 and q.n[0] >= 48 and f6 is None:
            f6 = q
        elif len(q.n) < 3 and q.sites * q.n_steps * (2 * q.F + 1) <= 1.5e8 and len(rest) < 7:
            rest.append(q)
    assert f6 is not None and len(rest) == 7
    return sorted([f6] + rest, key=lambda q: q.arrival)


def _oracle_request(q):
    """The oracle's result for a served request: find_ground_state (ITP) or n_steps
    split_step! (real time) from init_psi(state) in the harmonic trap."""
    from oracle import spinor_oracle as O
    og = O.make_grid(q.n, q.L)
    itp = q.kind == "itp"
    ows = O.make_workspace(og, q.F, O.init_psi(og, q.F, q.state), q.dt, itp,
                           omega=(q.omega,) * len(q.n), c0=q.c0, c1=q.c1, n_steps=q.n_steps)
    if itp:
        r = O.find_ground_state(ows, q.n_steps, q.tol)
        return r["energy"], r["steps"], r["converged"], ows.psi
    for _ in range(q.n_steps):
        O.split_step(ows)
    return O.total_energy(ows), q.n_steps, False, ows.psi


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_serving_stream_slice(precision):
    """Eight requests of the serving configuration's own stream, arriving on the wall clock:
    each result equals the same request served alone, bit for bit, including a 3-D spin-6
    request; in FP64 the 1-D and 2-D requests also equal the oracle (energy 1e-12, psi 1e-10,
    the same convergence step)."""
    from oracle import spinor_oracle as O
    reqs = _stream_slice(precision)
    t0 = reqs[0].arrival
    for q in reqs:
        q.arrival -= t0
    srv = Server(slots_per_pool=4, chunk_ms=5.0, keep_psi=True)
    res, stats = srv.run(reqs, realtime=True)
    print(stats)
    by = {r.rid: r for r in res}
    assert sorted(by) == sorted(q.rid for q in reqs)
    for q in reqs:
        m = by[q.rid]
        assert m.error is None and math.isfinite(m.energy)
        assert m.finished >= m.admitted >= m.arrival - 1e-9
        a = solve_alone(q)
        assert (m.steps, m.converged) == (a.steps, a.converged)
        assert m.energy == a.energy and np.array_equal(m.psi, a.psi), q.rid
        if precision == "fp64" and len(q.n) < 3:
            e, steps, conv, psi = _oracle_request(q)
            assert abs(m.energy - e) <= 1e-12 * abs(e), (q.rid, m.energy, e)
            if q.kind == "itp":
                assert (m.steps, m.converged) == (steps, conv)
            assert O.relative_l2(m.psi, psi) <= 1e-10
