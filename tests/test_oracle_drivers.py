"""Pins of the oracle's drivers (CPU only): run_simulation_yoshida!, the Thomas-Fermi
initial state, classify_phase, find_ground_state_multistart and scan_continuation.

Each check is fixed by something other than the driver's own code: a limit in which the
controller's decisions are known in advance (fixed dt, commuting operators, the rejection
floor), an exact final state, a textbook closed form (Thomas-Fermi chemical potential,
uniform box ground states), or an identity between independent quantities (pair-amplitude
weights against the spin density).  The reference tests restated are named per test.
"""
import math

import numpy as np
import pytest

from oracle import spinor_oracle as O


def _psi(og, D, seed, width=None):
    r = np.random.default_rng(seed)
    psi = r.normal(size=(D,) + og.shape) + 1j * r.normal(size=(D,) + og.shape)
    if width is not None:
        psi = psi * O.gaussian(og, (width,) * og.ndim)[None]
    return psi / math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)


# --- run_simulation_yoshida! (yoshida.jl:11-98) ---------------------------------------
def test_yoshida_driver_fixed_dt_limit():
    """dt_min = dt_init = dt_max: every attempt has dt_step = dt (the save-point clamp is
    undone by the dt_min floor) and none is rejected (dt_step > 1.01 dt_min never holds),
    so the driver is repeated Yoshida steps; the last step overshoots t_end = 5.5 dt."""
    og = O.make_grid((32,), (10.0,))
    psi0 = _psi(og, 3, 1, width=1.5)
    kw = dict(omega=(1.0,), c0=5.0, c1=-0.3, p=0.2, q=0.05)
    dt = 0.01
    ws = O.make_workspace(og, 1, psi0, dt, False, **kw)
    r = O.run_simulation_yoshida(ws, dt_init=dt, dt_min=dt, dt_max=dt, tol=1e-14, t_end=5.5 * dt,
                                 save_interval=2 * dt)
    assert (r["n_accepted"], r["n_rejected"], r["final_dt"]) == (6, 0, dt)
    ref = O.make_workspace(og, 1, psi0, dt, False, **kw)
    for _ in range(6):
        O.yoshida_core(ref, dt)
    assert np.max(np.abs(ws.psi - ref.psi)) <= 1e-14
    assert np.allclose(r["result"]["times"], [0.0, 0.02, 0.04, 0.06], rtol=0, atol=1e-15)
    assert ws.t == pytest.approx(0.06, abs=1e-15) and ws.step == 6


def test_yoshida_driver_commuting_operators():
    """No interaction, no trap, uniform Zeeman field: the kinetic and potential steps commute,
    so S2 and S4 are both exact and the error estimate is round-off, which makes the accept
    factor its cap 3.  dt: 1e-3, 3e-3, 9e-3, then the 1e-2 cap; the step that would cross
    the save point 0.02 is clamped onto it (and leaves dt alone), the run lands exactly on
    t_end; the final state is the exact propagator."""
    og = O.make_grid((32, 16), (12.0, 8.0))
    psi0 = _psi(og, 3, 2, width=1.5)
    p, q = 0.3, 0.1
    ws = O.make_workspace(og, 1, psi0, 1e-3, False, p=p, q=q)
    r = O.run_simulation_yoshida(ws, dt_init=1e-3, dt_min=1e-5, dt_max=1e-2, tol=1e-3,
                                 t_end=0.05, save_interval=0.02)
    assert (r["n_accepted"], r["n_rejected"]) == (7, 0)
    assert r["final_dt"] == pytest.approx(1e-2, rel=1e-12)
    assert np.allclose(r["result"]["times"], [0.0, 0.02, 0.04, 0.05], rtol=0, atol=1e-14)
    T = ws.t
    assert T == pytest.approx(0.05, abs=1e-14)
    E = np.array([-p * m + q * m * m for m in (1, 0, -1)])
    ph = np.fft.fftn(psi0, axes=(1, 2)) * np.exp(-0.5j * og.k2 * T)[None]
    exact = np.fft.ifftn(ph, axes=(1, 2)) * np.exp(-1j * E * T)[:, None, None]
    assert np.max(np.abs(ws.psi - exact)) <= 1e-12


def test_yoshida_driver_rejection_floor():
    """tol far below any error: every attempt above 1.01 dt_min is rejected with the floor
    factor 0.3 (1e-3, 3e-4, 9e-5, 2.7e-5, then dt_min), the step at dt_min is accepted, and the
    accept factor (far below 1) keeps dt at dt_min: 4 rejections, then 1e-5 steps to t_end."""
    og = O.make_grid((32,), (10.0,))
    psi0 = _psi(og, 3, 3, width=1.5)
    ws = O.make_workspace(og, 1, psi0, 1e-3, False, omega=(1.0,), c0=5.0, c1=-0.3)
    # t_end far away (a step shortened to t_end counts as clamped and is never rejected);
    # the run stops after 5 accepted steps
    r = O.run_simulation_yoshida(ws, dt_init=1e-3, dt_min=1e-5, dt_max=1e-2, tol=1e-40,
                                 t_end=1.0, save_interval=1.0, max_accepted=5)
    assert (r["n_accepted"], r["n_rejected"]) == (5, 4)
    assert r["final_dt"] == 1e-5
    assert ws.t == pytest.approx(5e-5, rel=1e-12)


def test_yoshida_driver_reference_adaptive_run():
    """test_split_step.jl:190-208: 1-D spin-1, 64 points, box 20, c0 = 1, Zeeman p = 0.1,
    ferromagnetic start; adaptive (0.005, 1e-4, 0.05, tol 0.01) to t = 0.5 saving every 0.1:
    steps accepted, at least two snapshots, norm conserved to 1e-6."""
    og = O.make_grid((64,), (20.0,))
    ws = O.make_workspace(og, 1, O.init_psi(og, 1, "ferromagnetic"), 0.01, False, c0=1.0, p=0.1)
    r = O.run_simulation_yoshida(ws, dt_init=0.005, dt_min=1e-4, dt_max=0.05, tol=0.01,
                                 t_end=0.5, save_interval=0.1)
    assert r["n_accepted"] > 0
    norms = r["result"]["norms"]
    assert len(r["result"]["times"]) >= 2
    assert abs(norms[-1] - norms[0]) / norms[0] < 1e-6


# --- Thomas-Fermi state (thomas_fermi.jl:15-77) ----------------------------------------
def test_thomas_fermi_1d_chemical_potential():
    """test_physics_level2.jl:6-47: harmonic trap omega = 1, c0 = 200,
    mu_TF = (3 c0 / (4 sqrt 2))^(2/3); the profile is (mu - x^2/2)/c0 inside R = sqrt(2 mu)."""
    c0 = 200.0
    og = O.make_grid((4096,), (40.0,))
    V = O.harmonic_potential(og, (1.0,))
    psi = O.thomas_fermi_psi(og, 1, V, c0)
    assert np.all(psi[0] == 0) and np.all(psi[2] == 0)
    n = np.abs(psi[1]) ** 2
    assert np.sum(n) * og.dV == pytest.approx(1.0, rel=1e-12)
    mu = (3 * c0 / (4 * math.sqrt(2))) ** (2 / 3)
    i0 = int(np.argmax(n))
    assert V[i0] + c0 * n[i0] == pytest.approx(mu, rel=1e-4)
    x = og.x[0]
    n_tf = np.maximum(0.0, (mu - x ** 2 / 2) / c0)
    assert np.sum(np.abs(n - n_tf)) * og.dV < 1e-3
    R = math.sqrt(2 * mu)
    assert np.max(np.abs(x[n > 0])) == pytest.approx(R, abs=2 * og.dx[0])


def test_thomas_fermi_3d_chemical_potential():
    """Isotropic 3-D harmonic trap: mu_TF = (15 c0 / (4 pi))^(2/5) / 2; spin-2 polar
    placement (only m = 0 populated)."""
    c0 = 2000.0
    og = O.make_grid((64, 64, 64), (16.0, 16.0, 16.0))
    V = O.harmonic_potential(og, (1.0, 1.0, 1.0))
    psi = O.thomas_fermi_psi(og, 2, V, c0)
    assert all(np.all(psi[c] == 0) for c in (0, 1, 3, 4))
    n = np.abs(psi[2]) ** 2
    mu = 0.5 * (15 * c0 / (4 * math.pi)) ** 0.4
    inside = n > 0
    # V + c0 n is the same mu at every populated site
    assert np.ptp((V + c0 * n)[inside]) <= 1e-9 * mu
    assert float(np.max((V + c0 * n)[inside])) == pytest.approx(mu, rel=2e-2)


# --- classify_phase (diagnostics.jl:166-212) --------------------------------------------
def _state(og, amps):
    g = O.gaussian(og, (1.5,) * og.ndim)
    psi = np.array([a * g for a in amps], dtype=complex)
    return psi / math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)


@pytest.mark.parametrize("F,amps,phase", [
    (1, (1, 0, 0), "ferromagnetic"),
    (1, (0, 0, 1), "ferromagnetic"),
    (1, (0, 1, 0), "polar"),
    (1, (1, 0, 1), "polar"),                      # polar state rotated: A00 = n / sqrt 3
    (2, (0, 0, 1, 0, 0), "nematic"),
    (2, (0.5, 0, 1j / math.sqrt(2), 0, 0.5), "mixed"),   # cyclic: w_4 / sum w = 3/7
    (6, (1,) + (0,) * 12, "ferromagnetic"),
])
def test_classify_phase_known_states(F, amps, phase):
    og = O.make_grid((16, 16), (8.0, 8.0))
    r = O.classify_phase(_state(og, amps), F, og)
    assert r["phase"] == phase
    if phase == "ferromagnetic":
        assert r["spin_order"] == pytest.approx(1.0, rel=1e-12)
    if phase in ("polar", "nematic"):
        assert r["nematic_order"] == pytest.approx(1.0, rel=1e-12)
    if F == 2 and phase == "mixed":
        assert r["channel_weights"][4] == pytest.approx(3 / 7, rel=1e-12)
        assert r["channel_weights"][0] == pytest.approx(0.0, abs=1e-14)
    assert O.classify_phase(np.zeros((2 * F + 1,) + og.shape, complex), F, og)["phase"] == "vacuum"


@pytest.mark.parametrize("F", [1, 2, 3])
def test_pair_weights_against_spin_density(F):
    """CG completeness: sum_S w_S = int n^2; and sum_S (S(S+1)/2 - F(F+1)) w_S = int |f|^2
    with f the spin density (observables.jl:61-88), for a random spinor field."""
    og = O.make_grid((8, 8), (4.0, 4.0))
    psi = _psi(og, 2 * F + 1, 5)
    dV = og.dV
    w = {S: sum(float(np.sum(np.abs(O.pair_amplitude(psi, F, S, M)) ** 2)) * dV
                for M in range(-S, S + 1)) for S in range(0, 2 * F + 1, 2)}
    n = np.sum(np.abs(psi) ** 2, axis=0)
    fx, fy, fz = O.spin_density(psi, F)
    assert sum(w.values()) == pytest.approx(float(np.sum(n ** 2)) * dV, rel=1e-12)
    lhs = sum((S * (S + 1) / 2 - F * (F + 1)) * v for S, v in w.items())
    assert lhs == pytest.approx(float(np.sum(fx ** 2 + fy ** 2 + fz ** 2)) * dV, rel=1e-11)


# --- find_ground_state_multistart / scan_continuation (ground_state.jl:239-378) ---------
L_BOX = 10.0
STATES = ("polar", "ferromagnetic", "uniform", "antiferromagnetic")


def _box_ws(og, c0, c1, psi0, n_steps, dt=0.05):
    return O.make_workspace(og, 1, psi0, dt, True, c0=c0, c1=c1, n_steps=n_steps)


def test_multistart_box_ground_state_closed_form():
    """Box without trap: the ground state is uniform, E = c0 / (2L) for the polar state and
    (c0 + c1) / (2L) for the ferromagnetic one (unit norm, n = 1/L).  With c1 < 0 the polar
    start stays polar (its m = +-1 components are zero) and the best start is ferromagnetic;
    with c1 > 0 the best is polar."""
    og = O.make_grid((16,), (L_BOX,))
    starts = [(s, 0, O.init_psi(og, 1, s)) for s in STATES]
    for c1, e_best, polar_best in ((-0.5, (10.0 - 0.5) / (2 * L_BOX), False),
                                   (0.5, 10.0 / (2 * L_BOX), True)):
        m = O.find_ground_state_multistart(lambda p0: _box_ws(og, 10.0, c1, p0, 1500), starts,
                                           1500, 1e-10)
        assert m["energy"] == pytest.approx(e_best, rel=1e-8)
        assert len(m["all_results"]) == 4
        e_pol = m["all_results"][0]["energy"]
        assert e_pol == pytest.approx(10.0 / (2 * L_BOX), rel=1e-8)
        assert (m["initial_state"] == "polar") == polar_best
        if not polar_best:
            assert m["energy"] < e_pol
        # first minimum: the best is the earliest start at the minimal energy
        es = [r["energy"] for r in m["all_results"]]
        assert m["energy"] == min(es) and es.index(min(es)) == STATES.index(m["initial_state"])


def test_continuation_fallback_and_energies():
    """Continuation over c0 = 10, 10.5, 20 with c1 = -0.5 from the polar state: the first two
    points continue the (stationary) polar state, E = c0 / (2L); the jump to c0 = 20 moves the
    energy by 90 % > 10 %, so the point is solved again by multistart, which finds the
    ferromagnetic ground state (c0 + c1) / (2L)."""
    og = O.make_grid((16,), (L_BOX,))
    starts = [(s, 0, O.init_psi(og, 1, s)) for s in STATES]
    out = O.scan_continuation([10.0, 10.5, 20.0], lambda v, p0, n=1500: _box_ws(og, v, -0.5, p0, n),
                              O.init_psi(og, 1, "polar"), starts, tol=1e-10,
                              n_steps_continuation=1500, n_steps_fresh=1500)
    assert [r["fallback"] for r in out] == [False, False, True]
    assert [r["phase"] for r in out] == ["polar", "polar", "ferromagnetic"]
    for r, e in zip(out, (10.0 / (2 * L_BOX), 10.5 / (2 * L_BOX), 19.5 / (2 * L_BOX))):
        assert r["energy"] == pytest.approx(e, rel=1e-8)
        assert r["converged"]


def test_continuation_reference_sweep():
    """test_continuation.jl:6-64 (grid 16, box 10, c0 = 10, dt = 0.01, tol 1e-6): five
    points keep their parameters and valid phases; nearby parameters (1.0, 1.01) give
    energies within 10 %."""
    og = O.make_grid((16,), (L_BOX,))
    starts = [(s, 0, O.init_psi(og, 1, s)) for s in STATES]

    def mk(v, p0, n=500):
        return O.make_workspace(og, 1, p0, 0.01, True, c0=10.0, c1=v, n_steps=n)

    out = O.scan_continuation([-2.0, -1.0, 0.0, 1.0, 2.0], mk, O.init_psi(og, 1, "polar"), starts,
                              tol=1e-6, n_steps_continuation=100, n_steps_fresh=500)
    assert [r["param"] for r in out] == [-2.0, -1.0, 0.0, 1.0, 2.0]
    assert all(r["phase"] in ("ferromagnetic", "polar", "nematic", "cyclic", "mixed", "vacuum")
               for r in out)
    near = O.scan_continuation([1.0, 1.01], mk, O.init_psi(og, 1, "polar"), starts, tol=1e-6,
                               n_steps_continuation=100, n_steps_fresh=500)
    assert abs(near[0]["energy"] - near[1]["energy"]) / abs(near[0]["energy"]) < 0.1
