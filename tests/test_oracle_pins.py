"""Pins of the FP64 oracle against closed forms, brute force and the reference's
own known-answer tests (CPU only).  Each test names the reference test it
restates; none of them re-types the oracle's own formula."""
import math

import numpy as np
import pytest
import scipy.linalg as sla
from sympy import Rational, sqrt as ssqrt
from sympy.physics.wigner import clebsch_gordan as s_cg

from oracle import spinor_oracle as O


def rng(seed=0):
    return np.random.default_rng(seed)


def rspinor(D, n, seed):
    r = rng(seed)
    return r.normal(size=(D, n)) + 1j * r.normal(size=(D, n))


# --- grid (grid.jl) --------------------------------------------------------------
def test_grid_cell_centred_and_fftfreq():
    g = O.make_grid((8, 6), (4.0, 3.0))
    assert np.allclose(g.x[0], -2 + 0.25 + 0.5 * np.arange(8), atol=1e-15)
    assert np.allclose(g.k[0], 2 * np.pi / 4 * np.array([0, 1, 2, 3, -4, -3, -2, -1]), atol=1e-14)
    assert np.allclose(g.k[1], 2 * np.pi / 3 * np.array([0, 1, 2, -3, -2, -1]), atol=1e-14)
    assert g.k2.shape == (6, 8)
    assert g.k2[2, 3] == pytest.approx(g.k[0][3] ** 2 + g.k[1][2] ** 2, rel=1e-15)
    with pytest.raises(ValueError):
        O.make_grid((7,), (1.0,))


# --- spin matrices (test_spin_matrices.jl) ---------------------------------------
@pytest.mark.parametrize("F", [1, 2, 3, 6])
def test_spin_algebra(F):
    s = O.spin_matrices(F)
    Fx, Fy, Fz = s["Fx"], s["Fy"], s["Fz"]
    assert np.allclose(Fx @ Fy - Fy @ Fx, 1j * Fz, atol=1e-12)
    assert np.allclose(Fy @ Fz - Fz @ Fy, 1j * Fx, atol=1e-12)
    cas = Fx @ Fx + Fy @ Fy + Fz @ Fz
    assert np.allclose(cas, F * (F + 1) * np.eye(2 * F + 1), atol=1e-11)
    assert np.allclose(s["fy_vals"], np.arange(-F, F + 1), atol=1e-12)


# --- Clebsch-Gordan / 6j (test_clebsch_gordan.jl) --------------------------------
def test_cg_known_values():
    assert O.cg(1, 1, -1, 0, 0) == pytest.approx(1 / math.sqrt(3), abs=1e-12)
    assert O.cg(1, 0, 0, 0, 0) == pytest.approx(-1 / math.sqrt(3), abs=1e-12)
    assert O.cg(1, 1, 1, 0, 0) == 0.0
    assert O.cg(2, 2, 2, 4, 4) == pytest.approx(1.0, abs=1e-12)
    for F in (1, 2, 3):
        for S in range(0, 2 * F + 1):
            for M in range(-S, S + 1):
                s = sum(O.cg(F, m1, M - m1, S, M) ** 2 for m1 in range(-F, F + 1)
                        if abs(M - m1) <= F)
                assert s == pytest.approx(1.0, abs=1e-10)


def test_6j_roundtrip_cn_gS():
    """c_k -> g_S -> back (test_clebsch_gordan.jl round-trip) and the F=1
    physical relation g0 = c0 - 2c1, g2 = c0 + c1 (interactions.jl:123-128)."""
    F = 2
    c = {0: 1.3, 2: -0.4, 4: 0.7}
    g = O.cn_to_gS(F, c)
    ev = [0, 2, 4]
    W = np.array([[(2 * k + 1) * O.wigner6j(F, F, k, F, F, S) for k in ev] for S in ev])
    back = np.linalg.solve(W, [g[S] for S in ev])
    assert np.allclose(back, [c[k] for k in ev], atol=1e-12)
    g1 = O.c0c1_to_gS(1, 3.0, 0.5)
    assert g1[0] == pytest.approx(3.0 - 2 * 0.5) and g1[2] == pytest.approx(3.5)


@pytest.mark.parametrize("F", [1, 2, 6])
def test_channel_identity(F):
    """sum_S g_S sum_M |A_SM|^2 = c0 n^2 + c1 |F|^2 for g_S = c0 + c1 (S(S+1)-2F(F+1))/2
    (test_interactions_constraint.jl:129-171) -- pins tensor_energy and the CG table."""
    c0, c1 = 2.7, -0.9
    ch = O.c0c1_to_gS(F, c0, c1)
    psi = rspinor(2 * F + 1, 5, F)
    E = O.tensor_energy(psi, F, ch, 1.0)           # = 1/2 sum_S g_S |A|^2
    n = np.sum(np.abs(psi) ** 2, axis=0)
    fx, fy, fz = O.spin_density(psi, F)
    rhs = 0.5 * np.sum(c0 * n ** 2 + c1 * (fx ** 2 + fy ** 2 + fz ** 2))
    assert E == pytest.approx(rhs, rel=1e-10)


# --- spin density (observables.jl:61-88) -----------------------------------------
@pytest.mark.parametrize("F", [1, 2, 6])
def test_spin_density_equals_expectation(F):
    s = O.spin_matrices(F)
    psi = rspinor(2 * F + 1, 4, 10 + F)
    fx, fy, fz = O.spin_density(psi, F)
    for i in range(4):
        v = psi[:, i]
        assert fx[i] == pytest.approx(np.real(np.vdot(v, s["Fx"] @ v)), rel=1e-12)
        assert fy[i] == pytest.approx(np.real(np.vdot(v, s["Fy"] @ v)), rel=1e-12)
        assert fz[i] == pytest.approx(np.real(np.vdot(v, s["Fz"] @ v)), rel=1e-12)


# --- Rodrigues / Euler rotations vs brute-force expm (test_spin_mixing.jl) --------
@pytest.mark.parametrize("itp", [False, True])
def test_rodrigues_vs_expm(itp):
    s = O.spin_matrices(1)
    psi = rspinor(3, 6, 3)
    c1, dt = 0.7, 0.13
    out = O.rodrigues_step(psi.copy(), c1, dt, itp)
    fx, fy, fz = O.spin_density(psi, 1)
    for i in range(6):
        H = c1 * (fx[i] * s["Fx"] + fy[i] * s["Fy"] + fz[i] * s["Fz"])
        U = sla.expm(-H * dt) if itp else sla.expm(-1j * H * dt)
        assert np.allclose(out[:, i], U @ psi[:, i], atol=1e-12)


@pytest.mark.parametrize("F", [2, 6])
@pytest.mark.parametrize("itp", [False, True])
def test_euler_vs_expm(F, itp):
    s = O.spin_matrices(F)
    D = 2 * F + 1
    psi = rspinor(D, 5, 7)
    r = rng(5)
    ph = r.normal(size=(3, 5))
    dt = 0.21
    out = O.euler_rotation(psi.copy(), ph[0], ph[1], ph[2], dt, F, itp)
    for i in range(5):
        H = ph[0, i] * s["Fx"] + ph[1, i] * s["Fy"] + ph[2, i] * s["Fz"]
        U = sla.expm(-H * dt) if itp else sla.expm(-1j * H * dt)
        assert np.allclose(out[:, i], U @ psi[:, i], atol=1e-11)
    # identity below the |phi||dt| < 1e-14 threshold (test_spin_mixing.jl Euler identity)
    z = np.zeros(5)
    same = O.euler_rotation(psi.copy(), z + 1e-16, z, z, 1.0, F, itp)
    assert np.array_equal(same, psi)


# --- tensor step (tensor_interaction.jl) -----------------------------------------
def _brute_h(F, ch, v):
    D = 2 * F + 1
    h = np.zeros((D, D), complex)
    for S, g in ch.items():
        for M in range(-S, S + 1):
            for m in range(-F, F + 1):
                mu = M - m
                if abs(mu) > F:
                    continue
                a = float(s_cg(F, F, S, m, mu, M))
                for mp in range(-F, F + 1):
                    nu = M - mp
                    if abs(nu) > F:
                        continue
                    b = float(s_cg(F, F, S, mp, nu, M))
                    h[F - m, F - mp] += g * a * b * np.conj(v[F - mu]) * v[F - nu]
    return h


@pytest.mark.parametrize("itp", [False, True])
def test_tensor_step_vs_expm(itp):
    F = 2
    ch = {0: 1.1, 2: -0.6, 4: 0.9}
    T = O.tensor_fold(F, ch)
    psi = rspinor(5, 3, 2)
    out = O.tensor_step(psi.copy(), T, 0.17, itp)
    for i in range(3):
        h = _brute_h(F, ch, psi[:, i])
        assert np.allclose(h, h.conj().T, atol=1e-12)
        U = sla.expm(-h * 0.17) if itp else sla.expm(-1j * h * 0.17)
        assert np.allclose(out[:, i], U @ psi[:, i], atol=1e-11)


def test_tensor_f1_mean_field_equals_contact_plus_spin():
    """F=1 tensor with g from (c0, c1): the mean-field action h psi equals
    (c0 n + c1 f.F) psi exactly, and the propagation agrees with diagonal(c0) +
    Rodrigues(c1) to 1e-4 at small dt (test_analytical_validation.jl:197-240)."""
    c0, c1 = 4.0, -1.3
    s = O.spin_matrices(1)
    T = O.tensor_fold(1, O.c0c1_to_gS(1, c0, c1))
    psi = rspinor(3, 7, 9)
    fx, fy, fz = O.spin_density(psi, 1)
    n = np.sum(np.abs(psi) ** 2, axis=0)
    for i in range(7):
        v = psi[:, i]
        h = np.einsum("abcd,c,d->ab", T, np.conj(v), v)
        rhs = c0 * n[i] * v + c1 * (fx[i] * s["Fx"] + fy[i] * s["Fy"] + fz[i] * s["Fz"]) @ v
        assert np.allclose(h @ v, rhs, atol=1e-12)
    psi = psi / np.sqrt(n.max())
    dt = 1e-3
    a = O.tensor_step(psi.copy(), T, dt, False)
    b = O.diagonal_step(psi.copy(), np.zeros(7), np.zeros(3), c0, 0.0, dt, False)
    b = O.rodrigues_step(b, c1, dt, False)
    assert np.max(np.abs(a - b)) < 1e-4


# --- nematic step (nematic.jl; restates test_nematic.jl) --------------------------
@pytest.mark.parametrize("F", [1, 2])
def test_nematic_itp_step_vs_expm(F):
    """ITP: each (m, -m) pair follows da/dtau = -V conj(b), db/dtau = -V conj(a) with V
    frozen at the pre-update A00; the oracle's cosh/sinh form must equal expm of that
    real-linear generator (nematic.jl:77-98)."""
    D = 2 * F + 1
    c2, dt = 0.8, 0.37
    psi = rspinor(D, 5, 11 + F)
    out = O.nematic_step(psi.copy(), F, c2, dt, True)
    for i in range(5):
        A = sum((1 if c % 2 == 0 else -1) * psi[c, i] * psi[D - 1 - c, i] for c in range(D))
        A /= math.sqrt(D)
        for c in range((D + 1) // 2):
            cp = D - 1 - c
            V = c2 * A / math.sqrt(D) * (1 if c % 2 == 0 else -1)
            B = np.array([[-V.real, -V.imag], [-V.imag, V.real]])   # -V conj(.) on (re, im)
            if c == cp:
                z = sla.expm(B * dt) @ np.array([psi[c, i].real, psi[c, i].imag])
                assert abs(out[c, i] - (z[0] + 1j * z[1])) < 1e-12
            else:
                M = np.zeros((4, 4))
                M[0:2, 2:4] = B
                M[2:4, 0:2] = B
                z = sla.expm(M * dt) @ np.array([psi[c, i].real, psi[c, i].imag,
                                                  psi[cp, i].real, psi[cp, i].imag])
                assert abs(out[c, i] - (z[0] + 1j * z[1])) < 1e-12
                assert abs(out[cp, i] - (z[2] + 1j * z[3])) < 1e-12


def test_nematic_singlet_amplitude_states():
    """test_nematic.jl:2-55: |A00|^2 = n^2/3 (F=1 polar), 0 (F=1 ferromagnetic),
    n^2/5 (F=2 polar)."""
    g = O.make_grid((64,), (20.0,))
    for F, state, den in [(1, "polar", 3), (1, "ferromagnetic", None), (2, "polar", 5)]:
        psi = O.init_psi(g, F, state)
        A2 = float(np.sum(np.abs(O.singlet_amplitude(psi, F)) ** 2))
        n = np.sum(np.abs(psi) ** 2, axis=0)
        if den is None:
            assert A2 < 1e-28
        else:
            assert abs(A2 - float(np.sum(n ** 2)) / den) < 1e-12 * A2


def test_nematic_step_reference_invariants():
    """test_nematic.jl:79-137,206-246: c2 = 0 and ferromagnetic states are left
    unchanged; one real-time step conserves the norm of uniform states (F=1, c2=50;
    F=2, c2=30) to 1e-12; m <-> -m exchange symmetry in real and imaginary time; an
    ITP step on the polar state keeps 0 < N < 2 N0."""
    g = O.make_grid((64,), (20.0,))
    psi = O.init_psi(g, 1, "uniform")
    assert np.array_equal(O.nematic_step(psi.copy(), 1, 0.0, 0.01, False), psi)
    fer = O.init_psi(g, 1, "ferromagnetic")
    assert np.max(np.abs(O.nematic_step(fer.copy(), 1, 50.0, 0.01, False) - fer)) < 1e-14
    for F, c2 in [(1, 50.0), (2, 30.0)]:
        psi = O.init_psi(g, F, "uniform")
        out = O.nematic_step(psi.copy(), F, c2, 0.01, False)
        n0, n1 = np.sum(np.abs(psi) ** 2), np.sum(np.abs(out) ** 2)
        assert abs(n1 - n0) / n0 < 1e-12
    v = np.array([0.5 * np.exp(0.7j), 0.3 + 0.2j, 0.4 * np.exp(-0.3j)]).reshape(3, 1)
    for itp in (False, True):
        a = O.nematic_step(v.copy(), 1, 50.0, 0.05, itp)
        b = O.nematic_step(v[::-1].copy(), 1, 50.0, 0.05, itp)
        assert abs(a[0, 0] - b[2, 0]) < 1e-14 and abs(a[2, 0] - b[0, 0]) < 1e-14
    pol = O.init_psi(g, 1, "polar")
    out = O.nematic_step(pol.copy(), 1, 50.0, 0.001, True)
    assert 0 < np.sum(np.abs(out) ** 2) < 2 * np.sum(np.abs(pol) ** 2)


@pytest.mark.parametrize("F", [1, 2])
def test_nematic_complex_state_norm_and_small_dt(F):
    """test_nematic.jl:138-204: 50 real-time steps of a complex single-site spinor keep
    the norm to 1e-6; at dt = 1e-6 the step equals psi_m - i V_m dt conj(psi_-m) to
    rtol 1e-4 (the GP equation, V_m = c2 (-1)^(F-m) A00 / sqrt(D))."""
    if F == 1:
        v = np.array([0.5 * np.exp(0.7j), 0.3 + 0.2j, 0.4 * np.exp(-0.3j)])
        c2 = 50.0
    else:
        v = np.array([0.4 * np.exp(0.5j), 0.3 * np.exp(-0.8j), 0.2 + 0.1j, 0.35 * np.exp(1.2j),
                      0.25 * np.exp(-0.3j)])
        c2 = 30.0
    D = 2 * F + 1
    psi = v.reshape(D, 1).copy()
    for _ in range(50):
        psi = O.nematic_step(psi, F, c2, 0.01, False)
    assert abs(np.sum(np.abs(psi) ** 2) - np.sum(np.abs(v) ** 2)) < 1e-6 * np.sum(np.abs(v) ** 2)
    dt = 1e-6
    out = O.nematic_step(v.reshape(D, 1).copy(), F, c2, dt, False)[:, 0]
    A = sum((1 if c % 2 == 0 else -1) * v[c] * v[D - 1 - c] for c in range(D)) / math.sqrt(D)
    for c in range(D):
        V = c2 * (1 if c % 2 == 0 else -1) * A / math.sqrt(D)
        exp = v[c] - 1j * V * dt * np.conj(v[D - 1 - c])
        assert abs(out[c] - exp) <= 1e-4 * abs(exp)


def test_nematic_energy_of_polar_and_ferro():
    """|A00|^2 = n^2/D for the F=2 polar state and 0 for the ferromagnetic one, so
    E_nematic = c2/(2D) sum n^2 dV and 0 (energy.jl:106-109, test_nematic.jl:57-77)."""
    g = O.make_grid((16,), (10.0,))
    for state, expect in [("polar", True), ("ferromagnetic", False)]:
        psi = O.init_psi(g, 2, state)
        ws = O.make_workspace(g, 2, psi, 0.01, True, c0=1.0, c_extra=(0.9,))
        E = O.energy_terms(ws)["tensor"]
        n = np.sum(np.abs(psi) ** 2, axis=0)
        ref = 0.9 / (2 * 5) * float(np.sum(n ** 2)) * g.dV if expect else 0.0
        assert abs(E - ref) < 1e-14


# --- loss step (losses.jl; restates test_losses.jl) -------------------------------
@pytest.mark.parametrize("F", [1, 2, 3, 6])
def test_dipolar_relaxation_rate_properties(F):
    """test_losses.jl:89-108: m = -F stable, mean rate = gamma_dr, rates >= 0; F=6: every
    other component has a rate > 0.01 at gamma_dr = 0.5."""
    r = O.dipolar_relaxation_rates(F, 1.0)
    assert len(r) == 2 * F + 1 and abs(r[-1]) < 1e-15
    assert abs(np.mean(r) - 1.0) < 1e-12 and np.all(r >= -1e-15)
    if F == 6:
        assert np.all(O.dipolar_relaxation_rates(6, 0.5)[:12] > 0.01)


def test_loss_step_reference_behaviour():
    """test_losses.jl:11-87,110-145,270-291: LossParams(0,0) keeps the norm; spin-1 uniform
    state: m=-1 untouched, m=+1 decays faster than m=0; spin-2: m=-2 untouched, m=+2
    fastest; the denser half loses a larger fraction; L3 alone is m-independent; 2-D step."""
    g = O.make_grid((64,), (20.0,))
    psi = O.init_psi(g, 1, "uniform")
    assert np.allclose(O.loss_step(psi, O.dipolar_relaxation_rates(1, 0.0), 0.0, 0.01), psi,
                       rtol=1e-14, atol=0)
    r1 = O.dipolar_relaxation_rates(1, 0.1)
    p1 = psi.copy()
    for _ in range(100):
        p1 = O.loss_step(p1, r1, 0.0, 0.01)
    b, a = [np.sum(np.abs(psi[c]) ** 2) for c in range(3)], [np.sum(np.abs(p1[c]) ** 2) for c in range(3)]
    assert abs(a[2] - b[2]) <= 1e-14 * b[2] and a[0] < b[0] and a[1] < b[1] and a[0] < a[1]
    psi2 = O.init_psi(g, 2, "uniform")
    r2 = O.dipolar_relaxation_rates(2, 0.1)
    p2 = psi2.copy()
    for _ in range(50):
        p2 = O.loss_step(p2, r2, 0.0, 0.01)
    b2 = [np.sum(np.abs(psi2[c]) ** 2) for c in range(5)]
    a2 = [np.sum(np.abs(p2[c]) ** 2) for c in range(5)]
    assert abs(a2[4] - b2[4]) <= 1e-14 * b2[4]
    assert all(a2[c] < b2[c] for c in range(4)) and a2[0] < a2[1] and a2[0] < a2[2]
    dens = np.zeros((3, 64), dtype=complex)
    dens[1, :32], dens[1, 32:] = 2.0, 0.5
    out = O.loss_step(dens, O.dipolar_relaxation_rates(1, 0.05), 0.0, 0.01)
    fl = 1 - np.sum(np.abs(out[1, :32]) ** 2) / np.sum(np.abs(dens[1, :32]) ** 2)
    fr = 1 - np.sum(np.abs(out[1, 32:]) ** 2) / np.sum(np.abs(dens[1, 32:]) ** 2)
    assert fl > fr
    p3 = psi.copy()
    for _ in range(50):
        p3 = O.loss_step(p3, np.zeros(3), 0.1, 0.01)
    pops = [np.sum(np.abs(p3[c]) ** 2) for c in range(3)]
    assert abs(pops[0] - pops[1]) < 1e-10 * pops[0] and abs(pops[1] - pops[2]) < 1e-10 * pops[1]
    q = np.full((3, 32, 32), 1 / math.sqrt(3 * 32 * 32), dtype=complex)
    qo = O.loss_step(q, r1, 0.0, 0.01)
    assert abs(np.sum(np.abs(qo[2]) ** 2) - np.sum(np.abs(q[2]) ** 2)) < 1e-14
    assert np.sum(np.abs(qo[0]) ** 2) < np.sum(np.abs(q[0]) ** 2)


def test_loss_step_single_site_closed_form():
    """One site, one step: psi_c * exp(-(gamma_c + L3) n dt / 2) with n the pre-update
    density; gamma_c of spin 1 from the CG values by hand: raw = (1/2 + 1/2, 1/2, 0)
    (|<1 1;2 -1|1 0>|^2 = 1/2, |<1 1;2 -2|1 -1>|^2 = 1/2... normalised to mean gamma)."""
    raw = np.array([float(s_cg(1, 2, 1, 1, -1, 0) ** 2 + s_cg(1, 2, 1, 1, -2, -1) ** 2),
                    float(s_cg(1, 2, 1, 0, -1, -1) ** 2), 0.0])
    gam = 0.3 * raw / raw.mean()
    assert np.allclose(O.dipolar_relaxation_rates(1, 0.3), gam, rtol=1e-14)
    v = np.array([0.6 + 0.1j, -0.3j, 0.5]).reshape(3, 1)
    n = float(np.sum(np.abs(v) ** 2))
    out = O.loss_step(v, gam, 0.02, 0.7)
    for c in range(3):
        assert abs(out[c, 0] - v[c, 0] * math.exp(-(gam[c] + 0.02) * n * 0.7 / 2)) < 1e-15


# --- constrained-Mz and adaptive-dt ITP (ground_state.jl) ------------------------
def test_constrained_normalisation_closed_form():
    """Spin-1 with an empty m=0 component: Mz(lambda) = tanh(2 lambda + ln(N+/N-)/2), so the
    Newton solution is lambda = (atanh(target) - ln(N+/N-)/2)/2; the result has unit norm and
    <F_z> = target, and target = current Mz leaves the state normalised but unchanged."""
    g = O.make_grid((32,), (10.0,))
    base = O.gaussian(g, (2.0,))
    psi = np.stack([0.8 * base, 0 * base, 0.5 * base]).astype(complex)
    Np, Nm = (float(np.sum(np.abs(psi[c]) ** 2)) for c in (0, 2))
    for target in (-0.4, 0.1, 0.7):
        out = O.normalize_constrained(psi, 1, g.dV, target)
        lam = (math.atanh(target) - 0.5 * math.log(Np / Nm)) / 2
        exp = O.normalize(np.stack([psi[0] * math.exp(lam), psi[1], psi[2] * math.exp(-lam)]), g.dV)
        assert np.max(np.abs(out - exp)) < 1e-12 * np.max(np.abs(exp))
        assert abs(O.total_norm(out, g.dV) - 1) < 1e-13
        assert abs(O.magnetization(out, 1, g.dV) - target) < 1e-12
    m0 = O.magnetization(O.normalize(psi, g.dV), 1, g.dV)
    out = O.normalize_constrained(psi, 1, g.dV, m0)
    assert np.max(np.abs(out - O.normalize(psi, g.dV))) < 1e-14


def test_adaptive_itp_reaches_fixed_dt_ground_state():
    """test_adaptive_dt.jl:6-25 (c0 = 10, c1 = -0.5 here): adaptive-dt ITP ends within 1 %
    of the fixed-dt energy; with c0 = c1 = 0 it reaches the harmonic E = 1/2."""
    g = O.make_grid((64,), (15.0,))
    psi0 = O.init_psi(g, 1, "polar")
    wf = O.make_workspace(g, 1, psi0, 0.001, True, omega=(1.0,), c0=10.0, c1=-0.5, n_steps=5000)
    rf = O.find_ground_state(wf, 5000, 1e-8)
    wa = O.make_workspace(g, 1, psi0, 0.001, True, omega=(1.0,), c0=10.0, c1=-0.5, n_steps=5000)
    ra = O.find_ground_state_adaptive(wa, 5000, 1e-8, 0.01)
    assert abs(rf["energy"] - ra["energy"]) / abs(rf["energy"]) < 0.01
    w0 = O.make_workspace(g, 1, psi0, 0.001, True, omega=(1.0,), n_steps=5000)
    r0 = O.find_ground_state_adaptive(w0, 5000, 1e-10, 0.01)
    assert abs(r0["energy"] - 0.5) < 1e-6


# --- kinetic step / diagonal step closed forms (test_physics_level0.jl) ----------
def test_kinetic_plane_wave_exact():
    g = O.make_grid((16, 8), (5.0, 3.0))
    kx, ky = g.k[0][3], g.k[1][-2]
    X = g.x[0][None, :]
    Y = g.x[1][:, None]
    pw = np.exp(1j * (kx * X + ky * Y))
    psi = np.stack([pw, 2 * pw, 0 * pw])
    dt = 0.37
    out = O.kinetic_step(psi, O.kinetic_phase(g, dt, False))
    assert np.allclose(out, psi * np.exp(-0.5j * (kx ** 2 + ky ** 2) * dt), atol=1e-12)
    out = O.kinetic_step(psi, O.kinetic_phase(g, dt, True))
    assert np.allclose(out, psi * np.exp(-0.5 * (kx ** 2 + ky ** 2) * dt), atol=1e-12)


def test_free_gaussian_spreading():
    """<x^2>(T) = s0^2/2 + T^2/(2 s0^2) (test_physics_level0.jl:7-48)."""
    s0, T, dt = 2.0, 3.0, 0.01
    g = O.make_grid((256,), (40.0,))
    psi = np.zeros((3, 256), complex)
    psi[1] = np.exp(-g.x[0] ** 2 / (2 * s0 ** 2))
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * g.dV)
    ws = O.make_workspace(g, 1, psi, dt, False, n_steps=300)
    for _ in range(300):
        O.split_step(ws)
    n = np.abs(ws.psi[1]) ** 2
    x = g.x[0]
    var = np.sum(x ** 2 * n) / np.sum(n)
    exact = s0 ** 2 / 2 + T ** 2 / (2 * s0 ** 2)
    assert abs(var - exact) / exact < 1e-3


def test_larmor_precession():
    """Linear Zeeman p: <Fx>(t) = cos(p t), <Fy> = -sin(p t) (test_physics_level3.jl:7-56)."""
    p, dt, nst = 1.0, 0.005, 200
    g = O.make_grid((32,), (10.0,))
    psi = np.zeros((3, 32), complex)
    psi[0], psi[1], psi[2] = 0.5, 1 / math.sqrt(2), 0.5
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * g.dV)
    ws = O.make_workspace(g, 1, psi, dt, False, p=p, n_steps=nst)
    for _ in range(nst):
        O.split_step(ws)
    fx, fy, fz = O.spin_density(ws.psi, 1)
    Nn = O.total_norm(ws.psi, g.dV)
    t = nst * dt
    assert np.sum(fx) * g.dV / Nn == pytest.approx(math.cos(p * t), abs=1e-9)
    assert np.sum(fy) * g.dV / Nn == pytest.approx(-math.sin(p * t), abs=1e-9)


def test_diagonal_step_closed_form():
    g = O.make_grid((8,), (4.0,))
    psi = rspinor(3, 8, 4)
    V = O.harmonic_potential(g, (1.3,))
    zee = O.zeeman_diag(1, 0.4, 0.2)
    c0, cl, dtf = 3.0, 0.5, 0.07
    n = np.sum(np.abs(psi) ** 2, axis=0)
    out = O.diagonal_step(psi.copy(), V, zee, c0, cl, dtf, False)
    for c in range(3):
        ph = np.exp(-1j * (V + c0 * n + cl * n ** 1.5 + zee[c]) * dtf)
        assert np.allclose(out[c], psi[c] * ph, atol=1e-13)
    out = O.diagonal_step(psi.copy(), V, zee, c0, cl, dtf, True)
    for c in range(3):
        f = np.exp(-(V + c0 * n + cl * n ** 1.5 + zee[c] - zee.min()) * dtf)
        assert np.allclose(out[c], psi[c] * f, atol=1e-13)


# --- energies (test_analytical_validation.jl:9-73) -------------------------------
@pytest.mark.parametrize("state,ec1", [("ferro", 1.0), ("polar", 0.0), ("afm", 0.0)])
def test_uniform_state_energies(state, ec1):
    g = O.make_grid((64,), (10.0,))
    n0 = 1.0 / 10.0
    psi = np.zeros((3, 64), complex)
    if state == "ferro":
        psi[0] = math.sqrt(n0)
    elif state == "polar":
        psi[1] = math.sqrt(n0)
    else:
        psi[0], psi[2] = math.sqrt(n0 / 2), -math.sqrt(n0 / 2)
    ws = O.make_workspace(g, 1, psi, 0.01, False, c0=100.0, c1=10.0)
    E = O.energy_terms(ws)
    assert E["c0"] == pytest.approx(0.5 * 100 * n0 ** 2 * 10, rel=1e-10)
    assert E["c1"] == pytest.approx(ec1 * 0.5 * 10 * n0 ** 2 * 10, abs=1e-12)
    assert E["kin"] == pytest.approx(0.0, abs=1e-20)


def test_ddi_energy_zero_uniform_polarised():
    g = O.make_grid((16, 16, 16), (10.0, 10.0, 10.0))
    n0 = 1.0 / 1000.0
    psi = np.zeros((3,) + g.shape, complex)
    psi[0] = math.sqrt(n0)
    ws = O.make_workspace(g, 1, psi, 0.01, False, ddi=True, cdd=1.0)
    assert abs(O.energy_terms(ws)["ddi"]) < 1e-12


def test_ddi_q_traceless_and_secular():
    """test_ddi.jl:2-18 (Q traceless off k=0) and :104-129 (secular identities)."""
    Q = O.ddi_q((8, 6, 4), (0.5, 0.6, 0.7))
    tr = Q["xx"] + Q["yy"] + Q["zz"]
    assert np.allclose(tr.ravel()[1:], 0.0, atol=1e-14) and tr.ravel()[0] == 0.0
    Qs = O.ddi_q((8, 6, 4), (0.5, 0.6, 0.7), secular=True)
    assert np.allclose(Qs["xx"], -Qs["zz"] / 2) and np.allclose(Qs["xy"], 0)


def _brute_phi(Fs, Q, cdd, n):
    """Phi from the explicit definition: Phi_a(r) = Re (1/N) sum_k w(kx) e^{i k.r}
    C sum_b Q_ab(k) F_b(k), the sum over the half spectrum kx = 0..nx/2 with
    w = 1 at kx = 0, nx/2 and 2 otherwise, and the imaginary part of the kx = 0
    and kx = nx/2 planes dropped after the y, z inverse sums (c2r semantics)."""
    nx, ny, nz = n
    N = nx * ny * nz
    X, Yv, Zv = np.arange(nx), np.arange(ny), np.arange(nz)
    Fk = []
    for F in Fs:  # brute-force forward DFT onto the half spectrum
        out = np.zeros((nz, ny, nx // 2 + 1), complex)
        for kz in range(nz):
            for ky in range(ny):
                for kx in range(nx // 2 + 1):
                    e = np.exp(-2j * np.pi * (kz * Zv[:, None, None] / nz + ky * Yv[None, :, None] / ny
                                              + kx * X[None, None, :] / nx))
                    out[kz, ky, kx] = np.sum(F * e)
        Fk.append(out)
    names = [["xx", "xy", "xz"], ["xy", "yy", "yz"], ["xz", "yz", "zz"]]
    res_all = []
    for a in range(3):
        Pk = cdd * sum(Q[names[a][b]] * Fk[b] for b in range(3))
        res = np.zeros((nz, ny, nx))
        for kx in range(nx // 2 + 1):
            # inverse over z, y (complex), then the c2r x-sum
            Zp = np.zeros((nz, ny), complex)
            for kz in range(nz):
                for ky in range(ny):
                    Zp += Pk[kz, ky, kx] * np.exp(2j * np.pi * (kz * Zv[:, None] / nz + ky * Yv[None, :] / ny))
            if kx == 0 or kx == nx // 2:
                res += np.real(Zp)[:, :, None] * np.cos(2 * np.pi * kx * X / nx)[None, None, :]
            else:
                res += 2 * np.real(Zp[:, :, None] * np.exp(2j * np.pi * kx * X / nx)[None, None, :])
        res_all.append(res / N)
    return res_all


def test_ddi_potential_brute_force():
    n, dx = (4, 6, 4), (0.9, 0.7, 1.1)
    r = rng(8)
    Fs = [r.normal(size=(n[2], n[1], n[0])) for _ in range(3)]
    Q = O.ddi_q(n, dx)
    phi = O.ddi_potential(*Fs, Q, 1.7)
    ref = _brute_phi(Fs, Q, 1.7, n)
    for a in range(3):
        assert np.allclose(phi[a], ref[a], atol=1e-12)


@pytest.mark.parametrize("secular", [False, True])
def test_ddi_potential_1d_closed_form(secular):
    """ddi.jl:51-80 with N = 1 (k_y = k_z = 0): off k = 0 the full tensor is
    diag(2/3, -1/3, -1/3), the secular one diag(1/6, 1/6, -1/3), and the k = 0 mode is
    dropped, so single Fourier modes come back scaled (Nyquist mode included)."""
    n, L = 16, 7.0
    x = np.arange(n) * (L / n)
    k1 = 2 * math.pi / L
    fx = np.cos(k1 * x)
    fy = 0.5 * np.sin(2 * k1 * x) + 0.25
    fz = 0.3 + np.cos(8 * k1 * x)            # the Nyquist mode
    Q = O.ddi_q((n,), (L / n,), secular=secular)
    px, py, pz = O.ddi_potential(fx, fy, fz, Q, 1.7)
    qa, qz = (1 / 6, -1 / 3) if secular else (2 / 3, -1 / 3)
    qb = qa if secular else -1 / 3
    assert np.allclose(px, 1.7 * qa * np.cos(k1 * x), atol=1e-13)
    assert np.allclose(py, 1.7 * qb * 0.5 * np.sin(2 * k1 * x), atol=1e-13)
    assert np.allclose(pz, 1.7 * qz * np.cos(8 * k1 * x), atol=1e-13)


def test_ddi_potential_padded_brute_force():
    """ddi_padded.jl:6-85: F in the first n entries of a (2n)^3 zero grid (same dx, so
    dk is halved), the explicit c2r convolution there, Phi cropped to the first n."""
    n, dx = (4, 2, 4), (0.9, 0.7, 1.1)
    r = rng(9)
    Fs = [r.normal(size=(n[2], n[1], n[0])) for _ in range(3)]
    pn = tuple(2 * v for v in n)
    Qp = O.ddi_q(pn, dx)
    phi = O.ddi_potential_padded(*Fs, Qp, 0.8)
    pads = []
    for F in Fs:
        P = np.zeros((pn[2], pn[1], pn[0]))
        P[:n[2], :n[1], :n[0]] = F
        pads.append(P)
    ref = _brute_phi(pads, Qp, 0.8, pn)
    for a in range(3):
        assert np.allclose(phi[a], ref[a][:n[2], :n[1], :n[0]], atol=1e-12)
    # the padded potential differs from the periodic one (no wrap-around images)
    per = O.ddi_potential(*Fs, O.ddi_q(n, dx), 0.8)
    assert np.max(np.abs(per[0] - phi[0])) > 1e-3


# --- ground state / integrators ---------------------------------------------------
def test_harmonic_ground_state_energy():
    """c0 = c1 = 0 ITP converges to E = 1/2 per dimension (textbook)."""
    g = O.make_grid((64, 64), (16.0, 16.0))
    psi = O.init_psi(g, 1, "polar")
    ws = O.make_workspace(g, 1, psi, 0.002, True, omega=(1.0, 1.0), n_steps=4000)
    r = O.find_ground_state(ws, 4000, 1e-12)
    assert r["energy"] == pytest.approx(1.0, abs=2e-6)


def test_strang_second_and_yoshida_fourth_order():
    """Halving dt: Strang error ratio > 3 (test_split_step.jl:49-78), Yoshida > 6
    (:159-188).  Reference solution = Yoshida at dt/16."""
    g = O.make_grid((32,), (12.0,))
    psi = O.init_psi(g, 1, "uniform") * np.exp(0.3j * g.x[0])[None]
    T = 0.4

    def run(core, dt):
        ws = O.make_workspace(g, 1, psi, dt, False, omega=(1.0,), c0=20.0, c1=-2.0, p=0.3, q=0.5)
        for _ in range(int(round(T / dt))):
            core(ws, dt)
        return ws.psi

    ref = run(O.yoshida_core, 0.4 / 128)
    es = [np.max(np.abs(run(O.strang_core, dt) - ref)) for dt in (0.04, 0.02)]
    ey = [np.max(np.abs(run(O.yoshida_core, dt) - ref)) for dt in (0.04, 0.02)]
    assert es[0] / es[1] > 3.0
    assert ey[0] / ey[1] > 6.0


def test_l2_change_phase_identity():
    """sum|dpsi|^2/sum|psi|^2 = 2(1 - cos phi) for a global phase (test_adaptive_dt.jl:103-106)."""
    psi = rspinor(3, 50, 1)
    phi = 0.3
    assert O.l2_change(psi * np.exp(1j * phi), psi) == pytest.approx(2 * (1 - math.cos(phi)), rel=1e-12)


def test_leapfrog_equals_strang_to_roundoff():
    """The merged V(dt) leapfrog reproduces split_step! up to the non-additivity of the
    symmetric inner splitting (simulation.jl:40-48): with c1 = 0 the V steps commute and
    the two schedules agree to roundoff."""
    g = O.make_grid((32, 16), (8.0, 6.0))
    psi = O.init_psi(g, 1, "uniform") * (1 + 0.1 * rng(3).normal(size=g.shape))[None]
    a = O.make_workspace(g, 1, psi, 0.01, False, omega=(1, 1), c0=5.0, n_steps=20, save_every=5)
    b = O.make_workspace(g, 1, psi, 0.01, False, omega=(1, 1), c0=5.0, n_steps=20, save_every=5)
    O.run_simulation(a)
    for _ in range(20):
        O.split_step(b)
    assert np.max(np.abs(a.psi - b.psi)) < 1e-12


def _trig_field(grid, modes, seed):
    """A real trigonometric polynomial of resolvable modes (|m| < n/2 per axis) and its exact
    gradient: f = sum_j a_j cos(k_j . r + p_j), grad f = -sum_j a_j k_j sin(k_j . r + p_j)."""
    r = rng(seed)
    nd = grid.ndim
    X = [O._axis_field(grid, d, grid.x[d]) for d in range(nd)]
    f = np.zeros(grid.shape)
    grad = [np.zeros(grid.shape) for _ in range(nd)]
    for _ in range(modes):
        kv = [2 * math.pi * int(r.integers(-(grid.n[d] // 2) + 1, grid.n[d] // 2)) / grid.L[d]
              for d in range(nd)]
        a, p = r.normal(), r.uniform(0, 2 * math.pi)
        ph = sum(kv[d] * X[d] for d in range(nd)) + p
        f = f + a * np.cos(ph)
        for d in range(nd):
            grad[d] = grad[d] - a * kv[d] * np.sin(ph)
    return f, grad


@pytest.mark.parametrize("n,L", [((16,), (5.0,)), ((16, 8), (6.0, 3.0)), ((8, 16, 4), (2.0, 7.0, 3.0))])
def test_fft_derivative_exact_on_trig_polynomials(n, L):
    """fft_utils.jl:4-44 is exact for band-limited fields: an anisotropic grid and random
    modes per axis fail any swapped axis, sign or scale."""
    g = O.make_grid(n, L)
    f, grad = _trig_field(g, 6, 11)
    scale = max(np.max(np.abs(v)) for v in grad)
    got = O.fft_gradient(f, g)
    for d in range(g.ndim):
        assert np.max(np.abs(got[d] - grad[d])) < 1e-12 * scale
        assert np.max(np.abs(O.fft_partial_derivative(f, g, d + 1) - grad[d])) < 1e-12 * scale


def test_fft_derivative_gaussian_spectral_accuracy():
    """d/dx exp(-x^2/2s^2) = -x/s^2 exp(..) to spectral accuracy for a well-resolved,
    well-contained Gaussian."""
    g = O.make_grid((128,), (20.0,))
    x = g.x[0]
    s = 1.0
    f = np.exp(-x ** 2 / (2 * s * s))
    assert np.max(np.abs(O.fft_partial_derivative(f, g, 1) + x / s ** 2 * f)) < 1e-12


def _adaptive_ws(c1=-0.5, n=64, L=20.0):
    g = O.make_grid((n,), (L,))
    psi = O.init_psi(g, 1, "uniform") * (1 + 0.05 * rng(7).normal(size=g.shape))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * g.dV)
    return g, psi


def test_adaptive_strang_fixed_dt_limits():
    """With dt_min = dt_max = dt_init nothing adapts: the step-change (FSAL) loop must equal
    split_step! (V(dt/2) K V(dt/2)) with the half potentials merged, and the Richardson loop
    two split steps of dt/2 each, to roundoff when the potential operators commute (c1 = 0:
    merging V(a) V(b) = V(a + b) is exact up to rounding)."""
    g, psi = _adaptive_ws()
    dt = 0.004
    for mode, sub in (("step_change", 1), ("richardson", 2)):
        a = O.make_workspace(g, 1, psi, dt, False, omega=(1.0,), c0=10.0, c1=0.0, q=0.2)
        out = O.run_simulation_adaptive(a, dt_init=dt, dt_min=dt, dt_max=dt, tol=1e-3,
                                        error_mode=mode, t_end=0.2, save_interval=0.1)
        assert out["n_rejected"] == 0 and out["n_accepted"] == 50
        b = O.make_workspace(g, 1, psi, dt / sub, False, omega=(1.0,), c0=10.0, c1=0.0, q=0.2)
        for _ in range(50 * sub):
            O.split_step(b)
        assert np.max(np.abs(a.psi - b.psi)) < 1e-11, mode
        assert out["result"]["times"] == pytest.approx([0.0, 0.1, 0.2])


@pytest.mark.parametrize("mode", ["step_change", "richardson"])
def test_adaptive_strang_reference_behaviour(mode):
    """test_adaptive_dt.jl:55-131: norm conserved to 1e-4 at every save, the final energy
    within 1 % of a fixed-dt leapfrog run, dt grows from a small initial value."""
    g, psi = _adaptive_ws()
    a = O.make_workspace(g, 1, psi, 0.001, False, omega=(1.0,), c0=10.0, c1=-0.5)
    out = O.run_simulation_adaptive(a, dt_init=0.001, dt_min=1e-5, dt_max=0.01, tol=1e-3,
                                    error_mode=mode, t_end=0.5, save_interval=0.1)
    rec = out["result"]
    assert out["n_accepted"] > 0 and len(rec["times"]) >= 3
    assert rec["times"][0] == 0.0 and rec["times"][-1] >= 0.5 - 0.01
    assert all(abs(nn - rec["norms"][0]) <= 1e-4 * rec["norms"][0] for nn in rec["norms"])
    f = O.make_workspace(g, 1, psi, 0.001, False, omega=(1.0,), c0=10.0, c1=-0.5, n_steps=500,
                         save_every=500)
    rf = O.run_simulation(f)
    assert abs(rf["energies"][-1] - rec["energies"][-1]) / abs(rf["energies"][-1]) < 0.01
    b = O.make_workspace(g, 1, psi, 0.001, False, omega=(1.0,), c0=10.0, c1=-0.5)
    out2 = O.run_simulation_adaptive(b, dt_init=1e-4, dt_min=1e-5, dt_max=0.01, tol=1e-3,
                                     error_mode=mode, t_end=0.5, save_interval=0.5)
    assert out2["final_dt"] > 1e-4


def test_adaptive_strang_rejects_and_shrinks():
    """A tolerance far below the first step's change forces rejections; each rejection
    shrinks dt by at least the 0.5 floor factor's bound and the accepted steps obey tol."""
    g, psi = _adaptive_ws()
    a = O.make_workspace(g, 1, psi, 0.01, False, omega=(1.0,), c0=10.0, c1=-0.5)
    out = O.run_simulation_adaptive(a, dt_init=0.01, dt_min=1e-5, dt_max=0.01, tol=1e-9,
                                    t_end=0.02, save_interval=0.02)
    assert out["n_rejected"] > 0 and out["final_dt"] < 0.01


# --- Raman coupling (test_raman.jl) ----------------------------------------------
def _m_plus_one(grid, D):
    psi = np.zeros((D,) + grid.shape, complex)
    psi[0] = 1.0 / math.sqrt(np.prod(grid.n) * grid.dV)
    return psi


@pytest.mark.parametrize("itp", [False, True])
def test_raman_spin1_rabi_closed_form(itp):
    """Spin 1 from m=+1 with delta = 0: U = e^{i a Fz} e^{-i beta Fx} e^{-i a Fz}, beta = Omega t,
    a = k.x, so psi = (cos^2(b/2), -i sin(b)/sqrt2 e^{-ia}, -sin^2(b/2) e^{-2ia}) psi_0 in real
    time and (cosh^2(b/2), -sinh(b)/sqrt2 e^{-ia}, sinh^2(b/2) e^{-2ia}) psi_0 in imaginary time.
    A factor 2 in Omega/2, the sign of k or a swapped F+/F- fails it."""
    g = O.make_grid((16,), (8.0,))
    k, Om, dt, n = 2 * math.pi / 4.0, 1.3, 0.01, 37
    psi = _m_plus_one(g, 3)
    a0 = psi[0].copy()
    for _ in range(n):
        psi = O.raman_step(psi, g, 1, (Om, 0.0, (k,)), dt, itp)
    b, ph = Om * dt * n, np.exp(-1j * k * np.asarray(g.x[0]))
    if itp:
        want = [np.cosh(b / 2) ** 2 * a0, -np.sinh(b) / math.sqrt(2) * ph * a0, np.sinh(b / 2) ** 2 * ph ** 2 * a0]
    else:
        want = [np.cos(b / 2) ** 2 * a0, -1j * np.sin(b) / math.sqrt(2) * ph * a0, -np.sin(b / 2) ** 2 * ph ** 2 * a0]
    for c in range(3):
        assert np.allclose(psi[c], want[c], atol=1e-13, rtol=0), c


def test_raman_detuning_only_phases_and_zero_coupling():
    """Omega = 0: psi_c *= e^{-i delta m dt} (m = F - c); Omega = delta = 0 leaves psi unchanged."""
    g = O.make_grid((8, 4), (4.0, 2.0))
    psi = rspinor(5, 32, 3).reshape(5, 4, 8)
    out = O.raman_step(psi, g, 2, (0.0, 0.7, (1.0, -2.0)), 0.05, False)
    for c in range(5):
        assert np.allclose(out[c], np.exp(-1j * 0.7 * (2 - c) * 0.05) * psi[c], atol=1e-14)
    assert np.allclose(O.raman_step(psi, g, 2, (0.0, 0.0, (0.0, 0.0)), 0.01, False), psi, atol=1e-14)


@pytest.mark.parametrize("F", [1, 2, 6])
def test_raman_step_matches_expm_and_keeps_norm(F):
    """Per-site scipy expm(-i H_R dt) with H_R assembled from the textbook spin matrices
    (F+ from sqrt(F(F+1) - m(m+1))); real-time norm conserved (test_raman.jl 'preserves norm')."""
    D = 2 * F + 1
    g = O.make_grid((4, 4), (3.0, 2.0))
    psi = rspinor(D, 16, 7).reshape(D, 4, 4)
    rc = (1.7, -0.4, (1.1, 0.6))
    out = O.raman_step(psi, g, F, rc, 0.03, False)
    m = np.array([F - c for c in range(D)], float)
    Fp = np.zeros((D, D))
    for c in range(1, D):   # <m+1| F+ |m>: row c-1 (m+1), column c (m)
        Fp[c - 1, c] = math.sqrt(F * (F + 1) - m[c] * (m[c] + 1))
    for iy in range(4):
        for ix in range(4):
            kr = 1.1 * g.x[0][ix] + 0.6 * g.x[1][iy]
            H = -0.4 * np.diag(m) + 0.85 * (np.exp(1j * kr) * Fp + np.exp(-1j * kr) * Fp.T)
            assert np.allclose(out[:, iy, ix], sla.expm(-1j * 0.03 * H) @ psi[:, iy, ix], atol=1e-13)
    assert np.sum(np.abs(out) ** 2) == pytest.approx(np.sum(np.abs(psi) ** 2), rel=1e-13)


def test_raman_energy_of_stretched_state_is_detuning():
    """<+F| H_R |+F> = delta F per site (the F+- terms have no diagonal), so E_raman = delta F * norm."""
    g = O.make_grid((16, 8), (8.0, 4.0))
    psi = _m_plus_one(g, 5)
    ws = O.make_workspace(g, 2, psi, 0.01, False, raman=(2.0, 0.3, (1.0, 0.5)))
    assert O.energy_terms(ws)["raman"] == pytest.approx(0.3 * 2 * 1.0, rel=1e-13)
