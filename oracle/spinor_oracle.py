"""FP64 CPU oracle for the SpinorBEC.jl split-step hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the `becgpu` package, the
HIP library) imports, links or calls this module.  Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` use it, and
only as the checker / reported CPU baseline.

It restates, with numpy (pocketfft for every FFT, LAPACK `eigh` for the
per-point Hermitian eigenproblems, sympy for exact Clebsch-Gordan / 6j), the
algorithms of the reference files cited function by function.  Array layout:
psi has numpy shape (D, nz, ny, nx) (C order), which is byte-identical to the
reference's column-major Julia array (nx, ny, nz, D): x fastest, component
slowest; component c <-> m = F - c.

Pinning (each function is checked in tests/test_oracle_*.py against closed
forms, invariants, brute force or reference-held golden data, never against
a retyped copy of its own formula):
  grid/k-space ............ numpy.fft.fftfreq + analytic cell-centred grid
  kinetic step ............ free-Gaussian spreading closed form; exact plane waves
  diagonal step ........... closed-form per-site phases; Larmor precession
  Rodrigues / Euler ....... scipy.linalg.expm(-i dt phi.F) brute force
  tensor step ............. scipy.linalg.expm of brute-force h; F=1 == spin mixing;
                            sum_S g_S |A_SM|^2 identity
  nematic step ............ ITP: scipy.linalg.expm of the real-linear pair ODE with A00
                            frozen; the reference's nematic tests (invariants, small dt)
  Raman step .............. spin-1 Rabi closed forms (real and imaginary time, with the
                            e^{-i k.x} phase of the transferred components); detuning-only
                            phases; scipy.linalg.expm of the per-site H_R; norm (test_raman.jl)
  loss step ............... the reference's loss tests (rates: m=-F stable, mean gamma_dr;
                            m-ordering of the decay; density dependence; L3 uniform)
  CG / 6j ................. sympy exact + the reference test's known values
  DDI ..................... direct O(N^2) k-sum convolution on small grids;
                            uniform polarised E_ddi = 0; padded variant: brute-force
                            linear (non-periodic) convolution on the doubled grid;
                            1-D: single Fourier modes scaled by the closed-form
                            diag(2/3, -1/3, -1/3) (secular diag(1/6, 1/6, -1/3))
  constrained Mz .......... closed form of the two-level Newton problem
  adaptive ITP ............ fixed-dt ITP limit; harmonic E = 1/2
  adaptive Strang ......... fixed-dt limits (step-change == split steps, Richardson == two
                            half steps) with commuting potentials; the reference's
                            run_simulation_adaptive! tests (norm, energy vs fixed dt, growth)
  FFT derivative .......... exact gradients of band-limited trigonometric fields on
                            anisotropic grids; Gaussian derivative closed form
  ITP / energy ............ uniform-state energies; c0=0 harmonic E = D/2;
                            3D F=6 golden density cuts (examples/eu151/density_64.json)
  Yoshida ................. coefficient closed form; 4th-order convergence ratio
  adaptive Yoshida driver . fixed-dt limit (dt_min = dt_max: repeated S4 steps, overshoot
                            of t_end, save points); commuting operators (err ~ 0: dt x3 per
                            step to dt_max, clamped steps landing on save points, exact
                            final state); the reference's adaptive test (norm, snapshots)
  Thomas-Fermi ............ 1-D mu_TF = (3 c0 / (4 sqrt 2))^(2/3), 3-D mu_TF =
                            (15 c0 / (4 pi))^(2/5) / 2 (harmonic trap, omega = 1)
  classify_phase .......... closed-form order parameters of the ferromagnetic, polar,
                            nematic and cyclic states
  multistart / continuation uniform ground states of the box: E = (c0 + min(c1, 0)) / (2L)
                            in 1-D; the best start and the continuation fallback follow
"""
from __future__ import annotations

import functools
import math
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------
# Grid  (grid.jl:1-35, 55-57)
# ---------------------------------------------------------------------------


@dataclass
class Grid:
    n: tuple          # Julia order (nx, ny, nz)
    L: tuple
    x: list           # per axis, Julia order
    dx: tuple
    k: list
    k2: np.ndarray    # numpy C-order shape (nz, ny, nx)

    @property
    def ndim(self):
        return len(self.n)

    @property
    def shape(self):
        return tuple(reversed(self.n))

    @property
    def dV(self):
        return float(np.prod(self.dx))


def make_grid(n, L) -> Grid:
    """grid.jl:1-23 (cell-centred x, fftfreq k) and :25-35 (k^2)."""
    n = tuple(int(v) for v in n)
    L = tuple(float(v) for v in L)
    for v in n:
        if v <= 0 or v % 2:
            raise ValueError("n_points must be positive even integers")  # types.jl:12
    xs, dxs = [], []
    for nd, Ld in zip(n, L):
        dx = Ld / nd
        xs.append(np.linspace(-Ld / 2 + dx / 2, Ld / 2 - dx / 2, nd))
        dxs.append(dx)
    # Julia fftfreq(n, n*dk) = j*dk; numpy fftfreq(n, d) = j/(n d) -> d = 1/(n dk)
    ks = [np.fft.fftfreq(nd, d=1.0 / (nd * 2 * np.pi / Ld)) for nd, Ld in zip(n, L)]
    shape = tuple(reversed(n))
    k2 = np.zeros(shape)
    nd_ = len(n)
    for d in range(nd_):
        sh = [1] * nd_
        sh[nd_ - 1 - d] = n[d]
        k2 = k2 + (ks[d] ** 2).reshape(sh)
    return Grid(n, L, xs, tuple(dxs), ks, k2)


def _axis_field(grid: Grid, d: int, vals):
    sh = [1] * grid.ndim
    sh[grid.ndim - 1 - d] = grid.n[d]
    return np.asarray(vals).reshape(sh)


# ---------------------------------------------------------------------------
# Spin matrices (spin_matrices.jl:1-50), Zeeman (zeeman.jl:15-19),
# potentials (potentials.jl:5-15)
# ---------------------------------------------------------------------------


@functools.lru_cache(maxsize=None)
def spin_matrices(F: int):
    D = 2 * F + 1
    m = np.array([F - c for c in range(D)], dtype=float)
    Fp = np.zeros((D, D), complex)
    for j in range(D):
        for i in range(D):
            if m[i] == m[j] + 1:
                Fp[i, j] = math.sqrt(F * (F + 1) - m[j] * (m[j] + 1))
    Fm = Fp.conj().T
    Fx = (Fp + Fm) / 2
    Fy = (Fp - Fm) / 2j
    Fz = np.diag(m).astype(complex)
    vals, vecs = np.linalg.eigh(Fy)
    assert np.all(np.abs(np.diff(vals) - 1.0) < 1e-10)
    return dict(F=F, D=D, m=m, Fx=Fx, Fy=Fy, Fz=Fz, Fp=Fp, Fm=Fm,
                fy_vals=vals, fy_vecs=vecs)


def zeeman_diag(F, p, q):
    """zeeman.jl:15-19: E_m = -p m + q m^2 for m = F..-F."""
    return np.array([-p * (F - c) + q * (F - c) ** 2 for c in range(2 * F + 1)])


def harmonic_potential(grid: Grid, omega):
    """potentials.jl:5-15: V = 1/2 sum_d omega_d^2 x_d^2."""
    V = np.zeros(grid.shape)
    for d in range(grid.ndim):
        V = V + omega[d] ** 2 * _axis_field(grid, d, grid.x[d] ** 2)
    return 0.5 * V


# ---------------------------------------------------------------------------
# Initial states (initialization.jl:1-79; thomas_fermi.jl:15-77)
# ---------------------------------------------------------------------------


def gaussian(grid: Grid, sigma):
    s = np.zeros(grid.shape)
    for d in range(grid.ndim):
        s = s + _axis_field(grid, d, grid.x[d] ** 2 / (2 * sigma[d] ** 2))
    return np.exp(-s)


def init_psi(grid: Grid, F: int, state="polar"):
    D = 2 * F + 1
    psi = np.zeros((D,) + grid.shape, complex)
    g = gaussian(grid, [L / 8 for L in grid.L])
    if state == "polar":
        psi[(D + 1) // 2 - 1] = g
    elif state == "ferromagnetic":
        psi[0] = g
    elif state == "uniform":
        for c in range(D):
            psi[c] = g / math.sqrt(D)
    elif state == "antiferromagnetic":
        for c in range(D):
            m = F - c
            psi[c] = (1.0 if (F - m) % 2 == 0 else -1.0) * g / math.sqrt(D)
    else:
        raise ValueError(f"Unknown initial state: {state}")
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * grid.dV)
    return psi


def _tf_norm(V, g, dV, mu):
    val = mu - V
    return np.sum(np.where(val > 0, val / g, 0.0)) * dV


def thomas_fermi_psi(grid: Grid, F: int, V, c0, N_target=1.0):
    """thomas_fermi.jl:15-77 (bisection for mu, polar component)."""
    dV = grid.dV
    Vmin, Vmax = V.min(), V.max()
    lo, hi = Vmin, Vmin + (Vmax - Vmin) * 2.0
    while _tf_norm(V, c0, dV, hi) < N_target and hi < Vmax * 100:
        hi *= 2.0
    mu = None
    for _ in range(200):
        mid = (lo + hi) / 2
        Nm = _tf_norm(V, c0, dV, mid)
        if abs(Nm - N_target) / N_target < 1e-12:
            mu = mid
            break
        if Nm < N_target:
            lo = mid
        else:
            hi = mid
    if mu is None:
        mu = (lo + hi) / 2
    n = np.maximum(0.0, (mu - V) / c0)
    D = 2 * F + 1
    psi = np.zeros((D,) + grid.shape, complex)
    psi[(D + 1) // 2 - 1] = np.sqrt(n)
    nrm = math.sqrt(np.sum(np.abs(psi) ** 2) * dV)
    if nrm > 0:
        psi = psi / nrm * math.sqrt(N_target)
    return psi


# ---------------------------------------------------------------------------
# Clebsch-Gordan / 6j (clebsch_gordan.jl) -- exact, via sympy
# ---------------------------------------------------------------------------


@functools.lru_cache(maxsize=None)
def cg(F, m1, m2, S, M):
    """<F m1; F m2 | S M> (clebsch_gordan.jl:69-73), exact via sympy."""
    if m1 + m2 != M or abs(M) > S:
        return 0.0
    from sympy.physics.wigner import clebsch_gordan
    return float(clebsch_gordan(F, F, S, m1, m2, M))


@functools.lru_cache(maxsize=None)
def wigner6j(a, b, c, d, e, f):
    from sympy.physics.wigner import wigner_6j
    return float(wigner_6j(a, b, c, d, e, f))


def cn_to_gS(F, c_dict):
    """clebsch_gordan.jl:117-141: g_S = sum_k (2k+1) {F F k; F F S} c_k."""
    return {S: sum((2 * k + 1) * wigner6j(F, F, k, F, F, S) * c_dict.get(k, 0.0)
                   for k in range(0, 2 * F + 1, 2))
            for S in range(0, 2 * F + 1, 2)}


def c0c1_to_gS(F, c0, c1):
    """interactions.jl:123-128."""
    return {S: c0 + c1 * (S * (S + 1) - 2 * F * (F + 1)) / 2 for S in range(0, 2 * F + 1, 2)}


def c_extra_to_delta_gS(F, c_extra):
    """interactions.jl:140-154 (c_extra[i] = c_{i+2} with 0-based i)."""
    c_dict = {}
    for idx, val in enumerate(c_extra, start=1):
        k = idx + 1
        if abs(val) > 1e-30 and k % 2 == 0 and k <= 2 * F:
            c_dict[k] = val
    if not c_dict:
        return {}
    return cn_to_gS(F, c_dict)


def tensor_channels(F, c0, c1, c_extra):
    """initialization.jl:164-185: activation rule and g_S = g_base + g_delta."""
    has_higher = any((idx + 1) % 2 == 0 and 4 <= idx + 1 <= 2 * F and abs(v) > 1e-30
                     for idx, v in enumerate(c_extra, start=1))
    if not has_higher:
        return None
    gb = c0c1_to_gS(F, c0, c1)
    gd = c_extra_to_delta_gS(F, c_extra)
    g = {S: gb.get(S, 0.0) + gd.get(S, 0.0) for S in range(0, 2 * F + 1, 2)}
    return {S: v for S, v in g.items() if abs(v) > 1e-30}


def tensor_fold(F, channels):
    """Dense T[m, m', mu, nu] = sum_S g_S CG(m,mu|S,M) CG(m',nu|S,M)
    (the HF-entry sum of tensor_interaction.jl:165-193 folded over S)."""
    D = 2 * F + 1
    T = np.zeros((D, D, D, D))
    for S, g in channels.items():
        for m in range(-F, F + 1):
            for mu in range(-F, F + 1):
                M = m + mu
                if abs(M) > S:
                    continue
                a = cg(F, m, mu, S, M)
                if abs(a) < 1e-15:
                    continue
                for mp in range(-F, F + 1):
                    nu = M - mp
                    if abs(nu) > F:
                        continue
                    b = cg(F, mp, nu, S, M)
                    if abs(b) < 1e-15:
                        continue
                    T[F - m, F - mp, F - mu, F - nu] += g * a * b
    return T


# ---------------------------------------------------------------------------
# Propagators
# ---------------------------------------------------------------------------


def density(psi):
    return np.sum(np.abs(psi) ** 2, axis=0)


def diagonal_step(psi, V, zee, c0, c_lhy, dtf, itp):
    """propagators.jl:82-119 (_diagonal_step_svec!): n from pre-update psi;
    ITP subtracts min Zeeman; RT uses cis."""
    n = density(psi)
    Vint = c0 * n + c_lhy * n * np.sqrt(n)
    if itp:
        zs = zee.min()
        base = np.exp(-(V + Vint) * dtf)
        for c in range(psi.shape[0]):
            psi[c] *= base * math.exp(-(zee[c] - zs) * dtf)
    else:
        base = np.exp(-1j * (V + Vint) * dtf)
        for c in range(psi.shape[0]):
            psi[c] *= base * np.exp(-1j * zee[c] * dtf)
    return psi


def spin_density(psi, F):
    """observables.jl:61-88: Fz = sum m|psi|^2, Fx + iFy = sum f+(m_c) conj(psi_{c-1}) psi_c."""
    D = 2 * F + 1
    m = np.array([F - c for c in range(D)], float)
    fz = np.einsum("c,c...->...", m, np.abs(psi) ** 2)
    fxy = np.zeros(psi.shape[1:], complex)
    for c in range(1, D):
        fp = math.sqrt(F * (F + 1) - m[c] * (m[c] + 1))
        fxy += fp * np.conj(psi[c - 1]) * psi[c]
    return fxy.real.copy(), fxy.imag.copy(), fz


def rodrigues_step(psi, c1, dtf, itp):
    """spin_mixing.jl:71-98 (F=1): U = I - i sin(th) nF + (cos th - 1) nF^2
    (ITP: I - sinh(th) nF + (cosh th - 1) nF^2), th = c1 |f| dt."""
    sm = spin_matrices(1)
    fx, fy, fz = spin_density(psi, 1)
    fmag = np.sqrt(fx ** 2 + fy ** 2 + fz ** 2)
    ok = fmag >= 1e-30
    safe = np.where(ok, fmag, 1.0)
    nF = (np.einsum("ab,...->ab...", sm["Fx"], fx / safe)
          + np.einsum("ab,...->ab...", sm["Fy"], fy / safe)
          + np.einsum("ab,...->ab...", sm["Fz"], fz / safe))
    nF2 = np.einsum("ab...,bc...->ac...", nF, nF)
    th = c1 * fmag * dtf
    eye = np.eye(3).reshape((3, 3) + (1,) * (psi.ndim - 1))
    if itp:
        U = eye - np.sinh(th) * nF + (np.cosh(th) - 1) * nF2
    else:
        U = eye - 1j * np.sin(th) * nF + (np.cos(th) - 1) * nF2
    out = np.einsum("ab...,b...->a...", U, psi)
    return np.where(ok[None], out, psi)


def euler_rotation(psi, phx, phy, phz, dt, F, itp):
    """spinor_utils.jl:180-264: exp(-i dt phi.F) psi = Rz(a) Ry(b) Dz(th) Ry(-b) Rz(-a) psi
    with b = acos(clamp(phi_z/|phi|)), a = atan2(phi_y, phi_x), th = |phi| dt,
    Ry via V_Fy diag(cis(+-b lambda)) V_Fy^dagger; identity if |phi||dt| < 1e-14."""
    sm = spin_matrices(F)
    D = 2 * F + 1
    V, lam = sm["fy_vecs"], sm["fy_vals"]
    m = sm["m"].reshape((D,) + (1,) * (psi.ndim - 1))
    lamr = lam.reshape((D,) + (1,) * (psi.ndim - 1))
    pm = np.sqrt(phx ** 2 + phy ** 2 + phz ** 2)
    skip = pm * abs(dt) < 1e-14
    safe = np.where(skip, 1.0, pm)
    beta = np.arccos(np.clip(phz / safe, -1.0, 1.0))
    alpha = np.arctan2(phy, phx)
    theta = pm * dt
    v = np.exp(1j * m * alpha) * psi                                  # Rz(-alpha)
    w = np.exp(1j * beta * lamr) * np.einsum("ij,j...->i...", V.conj().T, v)
    v = np.einsum("ij,j...->i...", V, w)                               # Ry(-beta)
    if itp:
        v = v * np.exp(-m * theta)
    else:
        v = v * np.exp(-1j * m * theta)                                # Dz(theta)
    w = np.exp(-1j * beta * lamr) * np.einsum("ij,j...->i...", V.conj().T, v)
    v = np.exp(-1j * m * alpha) * np.einsum("ij,j...->i...", V, w)     # Ry(beta), Rz(alpha)
    return np.where(skip[None], psi, v)


def spin_mixing_step(psi, F, c1, dtf, itp):
    """spin_mixing.jl:1-66."""
    if abs(c1) < 1e-30:
        return psi
    if F == 1:
        return rodrigues_step(psi, c1, dtf, itp)
    fx, fy, fz = spin_density(psi, F)
    return euler_rotation(psi, c1 * fx, c1 * fy, c1 * fz, dtf, F, itp)


def singlet_amplitude(psi, F):
    """observables.jl:94-111: A00 = sum_c (-1)^(F-m) psi_c psi_{D-1-c} / sqrt(D)."""
    D = 2 * F + 1
    A = np.zeros(psi.shape[1:], dtype=complex)
    for c in range(D):
        A += (1.0 if c % 2 == 0 else -1.0) * psi[c] * psi[D - 1 - c]
    return A / math.sqrt(D)


def nematic_step(psi, F, c2, dtf, itp):
    """nematic.jl:34-119: singlet-pair step with A00 from the pre-update spinor;
    each (m, -m) pair mixes with the conjugate of its partner."""
    if abs(c2) < 1e-30:
        return psi
    D = 2 * F + 1
    out = psi.copy()
    Vb = c2 * singlet_amplitude(psi, F) / math.sqrt(D)
    for c in range((D + 1) // 2):
        cp = D - 1 - c
        V = Vb * (1.0 if c % 2 == 0 else -1.0)
        a = np.abs(V)
        ok = a >= 1e-30
        ph = np.where(ok, V / np.where(ok, a, 1.0), 0.0)
        if itp:
            cc, ss = np.cosh(a * dtf), np.sinh(a * dtf)
            coef = ph * ss
        else:
            cc, ss = np.cos(a * dtf), np.sin(a * dtf)
            coef = 1j * ph * ss
        pm, pn = psi[c], psi[cp]
        if c == cp:
            out[c] = np.where(ok, cc * pm - coef * np.conj(pm), pm)
        else:
            out[c] = np.where(ok, cc * pm - coef * np.conj(pn), pm)
            out[cp] = np.where(ok, cc * pn - coef * np.conj(pm), pn)
    return out


def dipolar_relaxation_rates(F, gamma_dr):
    """losses.jl:60-82: gamma_m = gamma_dr sum_{q=-1,-2} CG(F,m;2,q|F,m+q)^2 / Z with Z the
    mean of the raw sums (sympy CG)."""
    from sympy.physics.wigner import clebsch_gordan as scg
    D = 2 * F + 1
    raw = []
    for c in range(D):
        m = F - c
        sacc = 0.0
        for q in (-1, -2):
            mp = m + q
            if abs(mp) > F:
                continue
            v = float(scg(F, 2, F, m, q, mp))
            sacc += v * v
        raw.append(sacc)
    tot = sum(raw)
    if tot < 1e-30:
        return np.zeros(D)
    Z = tot / D
    return np.array([gamma_dr * r / Z for r in raw])


def loss_step(psi, rates, L3, dt):
    """losses.jl:25-49: psi_c *= exp(-(gamma_c + L3) n dt / 2), n before the update;
    components with rate < 1e-30 are left alone."""
    out = psi.copy()
    n = density(psi)
    for c in range(psi.shape[0]):
        rate = rates[c] + L3
        if rate < 1e-30:
            continue
        out[c] = psi[c] * np.exp(-rate * n * dt / 2)
    return out


def tensor_step(psi, T, dtf, itp):
    """tensor_interaction.jl:195-240: h = sum T conj(psi_mu) psi_nu;
    psi <- V exp(-lambda dt) V^dagger psi (ITP) / V cis(-lambda dt) V^dagger psi;
    skipped where n < 1e-20."""
    D = psi.shape[0]
    sp = psi.reshape(D, -1)
    # h[p, a, b] = sum_{c,d} T[a, b, c, d] conj(psi_c) psi_d, as one (D^2 x D^2) @ (D^2 x N)
    # product over the sites
    pair = (np.conj(sp)[:, None, :] * sp[None, :, :]).reshape(D * D, -1)
    h = (T.reshape(D * D, D * D) @ pair).T.reshape(-1, D, D)
    lam, vec = np.linalg.eigh(h, UPLO="U")
    s = np.einsum("pjk,jp->pk", np.conj(vec), sp)
    ph = np.exp(-lam * dtf) if itp else np.exp(-1j * lam * dtf)
    out = np.einsum("pik,pk->ip", vec, ph * s)
    n = np.sum(np.abs(sp) ** 2, axis=0)
    out = np.where(n[None] < 1e-20, sp, out)
    return out.reshape(psi.shape)


def tensor_energy(psi, F, channels, dV):
    """tensor_interaction.jl:247-276: E = 1/2 sum_S g_S sum_M |A_SM|^2 dV."""
    E = 0.0
    for S, g in channels.items():
        for M in range(-S, S + 1):
            A = np.zeros(psi.shape[1:], complex)
            for m1 in range(-F, F + 1):
                m2 = M - m1
                if abs(m2) > F:
                    continue
                a = cg(F, m1, m2, S, M)
                if abs(a) < 1e-15:
                    continue
                A += a * psi[F - m1] * psi[F - m2]
            E += 0.5 * g * np.sum(np.abs(A) ** 2)
    return E * dV


# --- DDI (ddi.jl, ddi_padded.jl) ------------------------------------------


def ddi_q(n, dx, secular=False):
    """ddi.jl:7-82: Q_ab = k_a k_b/k^2 - delta_ab/3 on the rfft half grid
    (axis x from rfftfreq, Nyquist +; y,z from fftfreq, Nyquist -); Q(0) = 0."""
    nd = len(n)
    kx = np.arange(n[0] // 2 + 1) * (2 * np.pi / (n[0] * dx[0]))
    ky = np.fft.fftfreq(n[1], d=dx[1]) * 2 * np.pi if nd >= 2 else np.zeros(1)
    kz = np.fft.fftfreq(n[2], d=dx[2]) * 2 * np.pi if nd >= 3 else np.zeros(1)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    zero = k2 == 0.0
    inv = np.where(zero, 0.0, 1.0 / np.where(zero, 1.0, k2))
    if secular:
        qzz = np.where(zero, 0.0, KZ * KZ * inv - 1.0 / 3.0)
        Q = dict(xx=-qzz / 2, yy=-qzz / 2, zz=qzz, xy=0 * qzz, xz=0 * qzz, yz=0 * qzz)
    else:
        Q = dict(xx=np.where(zero, 0.0, KX * KX * inv - 1 / 3),
                 yy=np.where(zero, 0.0, KY * KY * inv - 1 / 3),
                 zz=np.where(zero, 0.0, KZ * KZ * inv - 1 / 3),
                 xy=KX * KY * inv, xz=KX * KZ * inv, yz=KY * KZ * inv)
    shp = tuple(reversed(n[:nd]))
    shp_h = shp[:-1] + (n[0] // 2 + 1,)
    return {k: v.reshape(shp_h) for k, v in Q.items()}


def ddi_potential(fx, fy, fz, Q, cdd):
    """ddi.jl:108-136 / :142-163: 3 rfft -> Phi_hat = C Q F_hat -> 3 irfft.
    numpy rfftn halves the LAST (x) axis, identical to the reference's axis 1;
    irfftn drops the imaginary parts of the x DC/Nyquist bins like FFTW c2r."""
    shp = fx.shape
    Fx, Fy, Fz = (np.fft.rfftn(a) for a in (fx, fy, fz))
    px = cdd * (Q["xx"] * Fx + Q["xy"] * Fy + Q["xz"] * Fz)
    py = cdd * (Q["xy"] * Fx + Q["yy"] * Fy + Q["yz"] * Fz)
    pz = cdd * (Q["xz"] * Fx + Q["yz"] * Fy + Q["zz"] * Fz)
    ax = tuple(range(len(shp)))
    return tuple(np.fft.irfftn(a, s=shp, axes=ax) for a in (px, py, pz))


def ddi_potential_padded(fx, fy, fz, Qp, cdd):
    """ddi_padded.jl:55-85: F into the first n entries of a 2x zero-padded grid,
    rfft convolution on the padded grid, crop the first n entries."""
    shp = fx.shape
    pshp = tuple(2 * s for s in shp)
    pads = []
    for a in (fx, fy, fz):
        P = np.zeros(pshp)
        P[tuple(slice(0, s) for s in shp)] = a
        pads.append(P)
    phis = ddi_potential(*pads, Qp, cdd)
    return tuple(p[tuple(slice(0, s) for s in shp)].copy() for p in phis)


# ---------------------------------------------------------------------------
# Workspace + integrator (types.jl:387-407, split_step.jl, propagators.jl)
# ---------------------------------------------------------------------------


@dataclass
class OracleWS:
    grid: Grid
    F: int
    psi: np.ndarray
    dt: float
    itp: bool
    V: np.ndarray
    p: float = 0.0
    q: float = 0.0
    c0: float = 0.0
    c1: float = 0.0
    c_lhy: float = 0.0
    c2: float = 0.0                # nematic coupling (c_extra[1]) when the tensor path is off
    loss: tuple = None             # (gamma_dr, L3) or None (LossParams)
    raman: tuple = None            # (Omega_R, delta, k_eff tuple) or None (RamanCoupling)
    zeeman_fn: object = None       # TimeDependentZeeman: t -> (p, q), or None
    channels: dict = None          # tensor path (g_S), or None
    T: np.ndarray = None
    cdd: float = 0.0
    ddi: bool = False
    Q: dict = None
    Qpad: dict = None
    normalize_every: int = 1
    save_every: int = 1
    n_steps: int = 1
    t: float = 0.0
    step: int = 0
    kphase: np.ndarray = field(default=None)

    @property
    def zee(self):
        """zeeman_at(ws.zeeman, ws.state.t) (zeeman.jl:20-21, split_step.jl:49-50)."""
        if self.zeeman_fn is not None:
            p, q = self.zeeman_fn(self.t)
            return zeeman_diag(self.F, p, q)
        return zeeman_diag(self.F, self.p, self.q)


def make_workspace(grid, F, psi, dt, itp, *, omega=None, V=None, p=0.0, q=0.0,
                   c0=0.0, c1=0.0, c_lhy=0.0, c_extra=(), ddi=False, cdd=0.0,
                   secular=False, padding=False, n_steps=1, normalize_every=None,
                   save_every=None, loss=None, raman=None, channels=None):
    """initialization.jl:86-191 (the subset on the hot path).  channels: an explicit
    tensor cache {S: g_S} (_make_tensor_cache_from_channels, tensor_interaction.jl:106-119)
    used instead of the activation rule; c0 and c1 are then left to the tensor step."""
    if V is None:
        V = harmonic_potential(grid, omega) if omega is not None else np.zeros(grid.shape)
    if channels is not None:
        ch = {S: float(g) for S, g in channels.items() if abs(g) > 1e-30}
    else:
        ch = tensor_channels(F, c0, c1, list(c_extra))
    T = None
    c2 = float(c_extra[0]) if len(c_extra) >= 1 else 0.0   # get_cn(ip, 2), types.jl:157-162
    if ch is not None:
        T = tensor_fold(F, ch)
        c0 = c1 = 0.0
        c2 = 0.0
    if normalize_every is None:
        normalize_every = 1 if itp else 0
    if save_every is None:
        save_every = max(1, n_steps // 100)           # types.jl:230
    ws = OracleWS(grid=grid, F=F, psi=np.array(psi, dtype=complex), dt=dt, itp=itp, V=V,
                  p=p, q=q, c0=c0, c1=c1, c_lhy=c_lhy, c2=c2, channels=ch, T=T, cdd=cdd, ddi=ddi,
                  normalize_every=normalize_every, save_every=save_every, n_steps=n_steps,
                  loss=None if loss is None else (float(loss[0]), float(loss[1])),
                  raman=None if raman is None else (float(raman[0]), float(raman[1]),
                                                    tuple(float(k) for k in raman[2])))
    if ddi:
        ws.Q = ddi_q(grid.n, grid.dx, secular)
        if padding:
            ws.Qpad = ddi_q(tuple(2 * v for v in grid.n), grid.dx, secular)
    ws.kphase = kinetic_phase(grid, dt, itp)
    return ws


def kinetic_phase(grid, dt, itp):
    """propagators.jl:1-7."""
    return np.exp(-0.5 * grid.k2 * dt) if itp else np.exp(-0.5j * grid.k2 * dt)


def kinetic_step(psi, phase):
    """propagators.jl:130-135: FFT over the spatial axes of every component,
    multiply by the phase, normalised inverse FFT."""
    ax = tuple(range(1, psi.ndim))
    return np.fft.ifftn(np.fft.fftn(psi, axes=ax) * phase[None], axes=ax)


def fft_partial_derivative(f, grid: Grid, dim):
    """fft_utils.jl:4-19: complex FFT of the real field, times i k_dim along the reference's
    axis dim (1 = x, the last numpy axis), normalised inverse FFT, real part."""
    fk = np.fft.fftn(np.asarray(f, dtype=complex))
    return np.real(np.fft.ifftn(1j * _axis_field(grid, dim - 1, grid.k[dim - 1]) * fk))


def fft_gradient(f, grid: Grid):
    """fft_utils.jl:25-44: one forward FFT, then the same multiply/inverse per axis."""
    fk = np.fft.fftn(np.asarray(f, dtype=complex))
    return tuple(np.real(np.fft.ifftn(1j * _axis_field(grid, d, grid.k[d]) * fk))
                 for d in range(grid.ndim))


def ddi_step(ws: OracleWS, dtf, itp):
    """ddi.jl:168-201 / ddi_padded.jl:90-124."""
    fx, fy, fz = spin_density(ws.psi, ws.F)
    if ws.Qpad is not None:
        phi = ddi_potential_padded(fx, fy, fz, ws.Qpad, ws.cdd)
    else:
        phi = ddi_potential(fx, fy, fz, ws.Q, ws.cdd)
    ws._last_phi = phi
    ws.psi = euler_rotation(ws.psi, *phi, dtf, ws.F, itp)


def raman_kr(grid: Grid, k_eff):
    """raman.jl:19: k_eff . r on the grid (numpy C-order shape)."""
    kr = np.zeros(grid.shape)
    for d in range(grid.ndim):
        kr = kr + _axis_field(grid, d, k_eff[d] * np.asarray(grid.x[d]))
    return kr


def raman_hamiltonians(grid: Grid, F, raman):
    """raman.jl:20-23: H_R = delta Fz + Omega_R/2 (e^{i k.r} F+ + e^{-i k.r} F-) per site,
    shape (N sites, D, D) in flat C order."""
    Omega, delta, keff = raman
    sm = spin_matrices(F)
    ph = np.exp(1j * raman_kr(grid, keff)).reshape(-1)
    return (delta * sm["Fz"][None]
            + (Omega / 2) * (ph[:, None, None] * sm["Fp"][None]
                             + np.conj(ph)[:, None, None] * sm["Fm"][None]))


def raman_step(psi, grid: Grid, F, raman, dtf, itp):
    """apply_raman_step! (raman.jl:8-31): psi <- U psi per site with
    U = _exp_i_hermitian(H_R, dt) (spinor_utils.jl:26-38): V diag(cis(-lam dt)) V^H,
    or exp(-lam dt) in imaginary time, from LAPACK eigh of the per-site H_R."""
    H = raman_hamiltonians(grid, F, raman)
    lam, V = np.linalg.eigh(H)
    e = np.exp(-lam * dtf) if itp else np.exp(-1j * lam * dtf)
    U = np.einsum("nij,nj,nkj->nik", V, e, V.conj())
    D = psi.shape[0]
    out = np.einsum("nij,nj->ni", U, psi.reshape(D, -1).T)
    return np.ascontiguousarray(out.T).reshape(psi.shape)


def raman_energy(psi, grid: Grid, F, raman, dV):
    """_raman_energy (energy.jl:123-135): sum_r Re(psi^H H_R psi) dV."""
    H = raman_hamiltonians(grid, F, raman)
    D = psi.shape[0]
    s = psi.reshape(D, -1).T
    return float(np.sum(np.real(np.einsum("ni,nij,nj->n", s.conj(), H, s)))) * dV


def half_potential_step(ws: OracleWS, dth, itp):
    """split_step.jl:48-133."""
    zee = ws.zee
    ws.psi = diagonal_step(ws.psi, ws.V, zee, ws.c0, ws.c_lhy, dth / 2, itp)
    if ws.T is not None:
        ws.psi = tensor_step(ws.psi, ws.T, dth / 2, itp)
    else:
        ws.psi = spin_mixing_step(ws.psi, ws.F, ws.c1, dth / 2, itp)
        ws.psi = nematic_step(ws.psi, ws.F, ws.c2, dth / 2, itp)
    if ws.raman is not None:
        ws.psi = raman_step(ws.psi, ws.grid, ws.F, ws.raman, dth / 2, itp)
    if ws.ddi:
        ddi_step(ws, dth, itp)
    if ws.raman is not None:
        ws.psi = raman_step(ws.psi, ws.grid, ws.F, ws.raman, dth / 2, itp)
    if ws.T is not None:
        ws.psi = tensor_step(ws.psi, ws.T, dth / 2, itp)
    else:
        ws.psi = nematic_step(ws.psi, ws.F, ws.c2, dth / 2, itp)
        ws.psi = spin_mixing_step(ws.psi, ws.F, ws.c1, dth / 2, itp)
    ws.psi = diagonal_step(ws.psi, ws.V, zee, ws.c0, ws.c_lhy, dth / 2, itp)


def normalize(psi, dV):
    """split_step.jl:185-195."""
    nrm2 = sum(np.sum(np.abs(psi[c]) ** 2) * dV for c in range(psi.shape[0]))
    return psi / math.sqrt(nrm2)


def apply_loss(ws: OracleWS, dt):
    """apply_loss_step! with the workspace's LossParams (no-op when absent or when both
    rates are below 1e-30, losses.jl:31)."""
    if ws.loss is None:
        return
    g, l3 = ws.loss
    if g < 1e-30 and l3 < 1e-30:
        return
    ws.psi = loss_step(ws.psi, dipolar_relaxation_rates(ws.F, g), l3, dt)


def split_step(ws: OracleWS):
    """split_step.jl:10-35."""
    half_potential_step(ws, ws.dt / 2, ws.itp)
    ws.psi = kinetic_step(ws.psi, ws.kphase)
    half_potential_step(ws, ws.dt / 2, ws.itp)
    if not ws.itp:
        apply_loss(ws, ws.dt)
    if not ws.itp:
        ws.t += ws.dt
    ws.step += 1
    if ws.itp and ws.normalize_every > 0 and ws.step % ws.normalize_every == 0:
        ws.psi = normalize(ws.psi, ws.grid.dV)


YOSHIDA_W1 = 1.0 / (2.0 - 2.0 ** (1 / 3))     # split_step.jl:139-140
YOSHIDA_W0 = 1.0 - 2.0 * YOSHIDA_W1


def strang_core(ws: OracleWS, dt):
    """split_step.jl:146-152 (real time, explicit dt)."""
    half_potential_step(ws, dt / 2, False)
    ws.psi = kinetic_step(ws.psi, kinetic_phase(ws.grid, dt, False))
    half_potential_step(ws, dt / 2, False)


def yoshida_core(ws: OracleWS, dt):
    """split_step.jl:161-183."""
    w1, w0 = YOSHIDA_W1, YOSHIDA_W0
    wm = (w1 + w0) / 2
    half_potential_step(ws, w1 * dt / 2, False)
    ws.psi = kinetic_step(ws.psi, kinetic_phase(ws.grid, w1 * dt, False))
    half_potential_step(ws, wm * dt, False)
    ws.psi = kinetic_step(ws.psi, kinetic_phase(ws.grid, w0 * dt, False))
    half_potential_step(ws, wm * dt, False)
    ws.psi = kinetic_step(ws.psi, kinetic_phase(ws.grid, w1 * dt, False))
    half_potential_step(ws, w1 * dt / 2, False)


def l2_change(psi_new, psi_old):
    """adaptive.jl:42-50."""
    return float(np.sum(np.abs(psi_new - psi_old) ** 2) / max(np.sum(np.abs(psi_old) ** 2), 1e-300))


# ---------------------------------------------------------------------------
# Energy and observables (energy.jl:6-121, observables.jl:1-33)
# ---------------------------------------------------------------------------


def energy_terms(ws: OracleWS):
    psi, g, dV = ws.psi, ws.grid, ws.grid.dV
    ax = tuple(range(1, psi.ndim))
    N = int(np.prod(g.n))
    ph = np.fft.fftn(psi, axes=ax)
    E = {}
    E["kin"] = 0.5 * sum(float(np.sum(g.k2 * np.abs(ph[c]) ** 2)) * dV / N for c in range(psi.shape[0]))
    E["trap"] = sum(float(np.sum(ws.V * np.abs(psi[c]) ** 2)) * dV for c in range(psi.shape[0]))
    zee = ws.zee
    E["zeeman"] = sum(zee[c] * float(np.sum(np.abs(psi[c]) ** 2)) * dV for c in range(psi.shape[0]))
    n = density(psi)
    if ws.T is None:
        E["c0"] = 0.5 * ws.c0 * float(np.sum(n ** 2)) * dV
        fx, fy, fz = spin_density(psi, ws.F)
        E["c1"] = 0.5 * ws.c1 * float(np.sum(fx ** 2 + fy ** 2 + fz ** 2)) * dV
        # energy.jl:43-46,106-109: the nematic term is reported in the tensor slot
        E["tensor"] = (0.5 * ws.c2 * float(np.sum(np.abs(singlet_amplitude(psi, ws.F)) ** 2)) * dV
                       if ws.c2 != 0.0 else 0.0)
    else:
        E["c0"] = 0.0
        E["c1"] = 0.0
        E["tensor"] = tensor_energy(psi, ws.F, ws.channels, dV)
    if ws.ddi:
        fx, fy, fz = spin_density(psi, ws.F)
        phi = ddi_potential(fx, fy, fz, ws.Q, ws.cdd)     # energy.jl:111-121: unpadded
        E["ddi"] = 0.5 * float(np.sum(phi[0] * fx + phi[1] * fy + phi[2] * fz)) * dV
    else:
        E["ddi"] = 0.0
    E["lhy"] = (2.0 / 5.0) * ws.c_lhy * float(np.sum(n * n * np.sqrt(n))) * dV if ws.c_lhy != 0 else 0.0
    E["raman"] = raman_energy(psi, g, ws.F, ws.raman, dV) if ws.raman is not None else 0.0
    return E


def total_energy(ws: OracleWS):
    E = energy_terms(ws)
    return (E["kin"] + E["trap"] + E["zeeman"] + E["c0"] + E["c1"] + E["ddi"]
            + E["lhy"] + E["tensor"] + E["raman"])


def total_norm(psi, dV):
    return float(np.sum(np.abs(psi) ** 2)) * dV


def magnetization(psi, F, dV):
    return sum((F - c) * float(np.sum(np.abs(psi[c]) ** 2)) * dV for c in range(psi.shape[0]))


# ---------------------------------------------------------------------------
# Drivers (ground_state.jl:80-163, simulation.jl:1-94, yoshida.jl:11-98)
# ---------------------------------------------------------------------------


def normalize_constrained(psi, F, dV, target_mz):
    """ground_state.jl:279-320: normalise, Newton on lambda (<= 20 iterations, clamp to
    [-10, 10]) so that the exp(lambda m)-weighted component norms have <F_z> = target,
    apply the weights, normalise again."""
    psi = normalize(psi, dV)
    D = 2 * F + 1
    lam = 0.0
    for _ in range(20):
        norms = np.array([float(np.sum(np.abs(psi[c]) ** 2)) * dV * math.exp(lam * (F - c)) ** 2
                          for c in range(D)])
        tot = float(norms.sum())
        if tot < 1e-30:
            break
        mz = sum((F - c) * norms[c] for c in range(D)) / tot
        if abs(mz - target_mz) < 1e-12:
            break
        dmz = sum(2 * (F - c) * ((F - c) - mz) * norms[c] / tot for c in range(D))
        if abs(dmz) < 1e-30:
            break
        lam -= (mz - target_mz) / dmz
        lam = min(max(lam, -10.0), 10.0)
    out = psi.copy()
    for c in range(D):
        out[c] = psi[c] * math.exp(lam * (F - c))
    return normalize(out, dV)


def find_ground_state(ws: OracleWS, n_steps, tol, target_mz=None):
    """ground_state.jl:80-163 after make_workspace: convergence every
    max(1, n_steps // 10) steps on |dE| and max|dpsi|/max|psi|.  With target_mz the
    workspace must have normalize_every = 0; the constrained normalisation runs at the
    start and after every step."""
    save_every = max(1, n_steps // 10)
    if target_mz is not None:
        ws.psi = normalize_constrained(ws.psi, ws.F, ws.grid.dV, target_mz)
    E_prev = total_energy(ws)
    psi_prev = ws.psi.copy()
    converged = False
    dE = dpsi = float("nan")
    steps = 0
    for step in range(1, n_steps + 1):
        split_step(ws)
        steps = step
        if step <= 10 and np.any(np.isnan(ws.psi)):
            raise ValueError(f"NaN detected in ITP at step {step}")
        if target_mz is not None:
            ws.psi = normalize_constrained(ws.psi, ws.F, ws.grid.dV, target_mz)
        if step % save_every == 0:
            E = total_energy(ws)
            dE = abs(E - E_prev)
            pmax = np.max(np.abs(ws.psi))
            dpsi = float(np.max(np.abs(ws.psi - psi_prev)) / pmax) if pmax > 0 else 0.0
            if dE < tol and dpsi < tol:
                converged = True
                break
            E_prev = E
            psi_prev = ws.psi.copy()
    return dict(converged=converged, energy=total_energy(ws), dE=dE, dpsi=dpsi, steps=steps)


def find_ground_state_adaptive(ws: OracleWS, n_steps, tol, dt_max):
    """ground_state.jl:172-228: check every max(1, n_steps // 100) steps; an energy rise
    (or NaN) restores psi and halves dt (floor 1e-8), otherwise convergence test and
    dt *= 1.1 up to dt_max.  ws: ITP workspace with normalize_every = 1."""
    check_every = max(1, n_steps // 100)
    E_prev = total_energy(ws)
    converged = False
    total = 0
    dE = dpsi = float("nan")
    while total < n_steps:
        backup = ws.psi.copy()
        for _ in range(check_every):
            split_step(ws)
        total += check_every
        E = total_energy(ws)
        if math.isnan(E) or E > E_prev:
            ws.psi = backup
            ws.dt = max(ws.dt * 0.5, 1e-8)
            ws.kphase = kinetic_phase(ws.grid, ws.dt, True)
        else:
            dE = abs(E - E_prev)
            pmax = np.max(np.abs(ws.psi))
            dpsi = float(np.max(np.abs(ws.psi - backup)) / pmax) if pmax > 0 else 0.0
            if dE < tol and dpsi < tol:
                converged = True
                break
            E_prev = E
            new_dt = min(ws.dt * 1.1, dt_max)
            if new_dt != ws.dt:
                ws.dt = new_dt
                ws.kphase = kinetic_phase(ws.grid, ws.dt, True)
    return dict(converged=converged, energy=total_energy(ws), dE=dE, dpsi=dpsi, steps=total,
                final_dt=ws.dt)


def find_ground_state_multistart(make_ws, starts, n_steps, tol):
    """ground_state.jl:239-271: one fresh find_ground_state per start, in order; best =
    the first minimum of the energies (Julia argmin).  starts: [(label, idx, psi0)] in the
    reference's order (each named state with idx 0; a :random entry expands to idx
    1..n_random at its place); make_ws(psi0) builds the ITP workspace of one solve."""
    results = []
    for label, idx, psi0 in starts:
        ws = make_ws(psi0)
        r = find_ground_state(ws, n_steps, tol)
        results.append(dict(initial_state=label, idx=idx, ws=ws, **r))
    best = results[0]
    for r in results[1:]:
        if r["energy"] < best["energy"]:
            best = r
    return dict(ws=best["ws"], converged=best["converged"], energy=best["energy"],
                initial_state=best["initial_state"], all_results=results)


def scan_continuation(param_values, make_ws, initial_psi, multistart_starts, *, tol,
                      energy_jump_threshold=0.1, n_steps_continuation=500, n_steps_fresh=5000):
    """ground_state.jl:328-378: the first point from initial_psi with n_steps_fresh steps,
    every later point from a copy of the previous ground state with n_steps_continuation
    steps; if the energy moved by more than energy_jump_threshold relative to the previous
    point, the point is solved again by find_ground_state_multistart (n_steps_fresh) over
    multistart_starts.  make_ws(val, psi0) builds the ITP workspace for parameter val.
    Each record also says whether the multistart fallback ran."""
    out = []
    prev_psi, prev_E = None, float("nan")
    for val in param_values:
        if prev_psi is not None:
            ws = make_ws(val, prev_psi.copy())
            r = find_ground_state(ws, n_steps_continuation, tol)
        else:
            ws = make_ws(val, np.array(initial_psi, dtype=complex))
            r = find_ground_state(ws, n_steps_fresh, tol)
        fallback = False
        if (not math.isnan(prev_E)
                and abs(r["energy"] - prev_E) / max(abs(prev_E), 1e-30) > energy_jump_threshold):
            m = find_ground_state_multistart(lambda p0, v=val: make_ws(v, p0), multistart_starts,
                                             n_steps_fresh, tol)
            ws, r, fallback = m["ws"], m, True
        ph = classify_phase(ws.psi, ws.F, ws.grid)
        out.append(dict(param=val, energy=r["energy"], converged=r["converged"],
                        phase=ph["phase"], psi=ws.psi.copy(), fallback=fallback))
        prev_psi, prev_E = ws.psi.copy(), r["energy"]
    return out


def pair_amplitude(psi, F, S, M):
    """observables.jl:118-143: A_SM = sum_{m1} CG(F m1; F M-m1 | S M) psi_m1 psi_{M-m1}
    (entries with |CG| < 1e-15 skipped, clebsch_gordan.jl:166-180)."""
    A = np.zeros(psi.shape[1:], complex)
    for m1 in range(-F, F + 1):
        m2 = M - m1
        if abs(m2) > F:
            continue
        a = cg(F, m1, m2, S, M)
        if abs(a) < 1e-15:
            continue
        A += a * psi[F - m1] * psi[F - m2]
    return A


def classify_phase(psi, F, grid: Grid):
    """diagnostics.jl:166-212 with pair_amplitude_spectrum (observables.jl:227-246):
    spin order sum|f|^2 dV / (F^2 sum n^2 dV); channel weights w_S = sum_M sum|A_SM|^2 dV;
    nematic order w_0 / (sum n^2 dV / D); phase ferromagnetic (spin order > 0.9), polar
    (F = 1) or nematic (nematic order > 0.9), cyclic (normalised w_2F > 0.5), else mixed;
    vacuum below a total norm of 1e-30."""
    D = 2 * F + 1
    dV = grid.dV
    n = density(psi)
    n_sum = float(np.sum(n)) * dV
    n_sq = float(np.sum(n ** 2)) * dV
    if n_sum < 1e-30:
        return dict(spin_order=0.0, nematic_order=0.0, channel_weights={}, phase="vacuum",
                    magnetization_density=0.0)
    fx, fy, fz = spin_density(psi, F)
    spin_order = float(np.sum(fx ** 2 + fy ** 2 + fz ** 2)) * dV / (float(F) ** 2 * n_sq)
    w = {S: sum(float(np.sum(np.abs(pair_amplitude(psi, F, S, M)) ** 2)) * dV
                for M in range(-S, S + 1)) for S in range(0, 2 * F + 1, 2)}
    tot = sum(w.values())
    cw = {S: (v / tot if tot > 0 else 0.0) for S, v in w.items()}
    nematic_order = w.get(0, 0.0) / (n_sq / D)
    mz = magnetization(psi, F, dV) / n_sum
    if spin_order > 0.9:
        phase = "ferromagnetic"
    elif nematic_order > 0.9:
        phase = "polar" if F == 1 else "nematic"
    elif cw.get(2 * F, 0.0) > 0.5:
        phase = "cyclic"
    else:
        phase = "mixed"
    return dict(spin_order=spin_order, nematic_order=nematic_order, channel_weights=cw,
                phase=phase, magnetization_density=mz)


def _snapshot(ws, rec):
    rec["times"].append(ws.t)
    rec["energies"].append(total_energy(ws))
    rec["norms"].append(total_norm(ws.psi, ws.grid.dV))
    rec["mags"].append(magnetization(ws.psi, ws.F, ws.grid.dV))


def run_simulation(ws: OracleWS, keep_psi=False):
    """simulation.jl:1-94: ITP -> split_step loop; RT -> leapfrog with merged
    V(dt) and splits at save points."""
    rec = dict(times=[], energies=[], norms=[], mags=[], psi=[])
    _snapshot(ws, rec)
    if ws.itp:
        for step in range(1, ws.n_steps + 1):
            split_step(ws)
            if step % ws.save_every == 0:
                _snapshot(ws, rec)
                if keep_psi:
                    rec["psi"].append(ws.psi.copy())
        return rec
    dt = ws.dt
    half_potential_step(ws, dt / 2, False)
    for step in range(1, ws.n_steps + 1):
        ws.psi = kinetic_step(ws.psi, ws.kphase)
        is_save = step % ws.save_every == 0
        is_last = step == ws.n_steps
        split = is_save or is_last
        half_potential_step(ws, dt / 2 if split else dt, False)
        apply_loss(ws, dt)                              # simulation.jl:74-76
        ws.t += dt
        ws.step += 1
        if is_save:
            _snapshot(ws, rec)
            if keep_psi:
                rec["psi"].append(ws.psi.copy())
        if split and not is_last:
            half_potential_step(ws, dt / 2, False)
    return rec


def run_simulation_yoshida(ws: OracleWS, *, dt_init=1e-3, dt_min=1e-5, dt_max=1e-2, tol=1e-3,
                           t_end, save_interval, max_accepted=None):
    """yoshida.jl:11-98 (AdaptiveDtParams defaults types.jl:343-361)."""
    dt = min(max(dt_init, dt_min), dt_max)
    rec = dict(times=[], energies=[], norms=[], mags=[], psi=[])
    _snapshot(ws, rec)
    next_save = ws.t + save_interval
    n_acc = n_rej = 0
    while ws.t < t_end - 1e-14:
        if max_accepted is not None and n_acc >= max_accepted:
            break
        dt_step = min(dt, t_end - ws.t)
        rem = next_save - ws.t
        if 1e-14 < rem < dt_step:
            dt_step = rem
        dt_step = max(dt_step, dt_min)
        clamped = dt_step < dt * 0.99
        psi_old = ws.psi.copy()
        strang_core(ws, dt_step)
        psi_s = ws.psi.copy()
        ws.psi = psi_old.copy()
        yoshida_core(ws, dt_step)
        err = l2_change(ws.psi, psi_s)
        if (not clamped) and err > tol and dt_step > dt_min * 1.01:
            ws.psi = psi_old
            dt = max(dt_step * max(0.3, 0.9 * (tol / err) ** 0.2), dt_min)
            n_rej += 1
            continue
        apply_loss(ws, dt_step)                         # yoshida.jl:64-66
        ws.t += dt_step
        ws.step += 1
        n_acc += 1
        if not clamped:
            fac = min(3.0, 0.9 * (tol / err) ** 0.2) if err > 1e-300 else 3.0
            dt = min(max(dt * fac, dt_min), dt_max)
        if ws.t >= next_save - 1e-14:
            _snapshot(ws, rec)
            next_save += save_interval
    if not rec["times"] or abs(rec["times"][-1] - ws.t) > 1e-12:
        _snapshot(ws, rec)
    return dict(result=rec, n_accepted=n_acc, n_rejected=n_rej, final_dt=dt)


def run_simulation_adaptive(ws: OracleWS, *, dt_init=1e-3, dt_min=1e-5, dt_max=1e-2, tol=1e-3,
                            error_mode="step_change", t_end, save_interval):
    """adaptive.jl:57-67 (run_simulation_adaptive!, real time): the step-change loop
    (adaptive.jl:69-194: V(dt/2) K(dt) with the closing half potential deferred and merged
    into the next step (FSAL), error = L2 change of the step, factors 0.9 sqrt(tol/err)
    clamped to [0.5, 2]) or the Richardson loop (adaptive.jl:203-296: one Strang step of dt
    against two of dt/2, factors 0.9 cbrt(tol/err))."""
    dt = min(max(dt_init, dt_min), dt_max)
    rec = dict(times=[], energies=[], norms=[], mags=[], psi=[])
    _snapshot(ws, rec)
    next_save = ws.t + save_interval
    n_acc = n_rej = 0
    fsal, fsal_dt = False, 0.0

    def hp(dth):
        half_potential_step(ws, dth, False)

    while ws.t < t_end - 1e-14:
        dt_step = min(dt, t_end - ws.t)
        rem = next_save - ws.t
        if 1e-14 < rem < dt_step:
            dt_step = rem
        dt_step = max(dt_step, dt_min)
        clamped = dt_step < dt * 0.99
        may_reject = (not clamped) and dt_step > dt_min * 1.01
        if error_mode == "richardson":
            saved = ws.psi.copy()
            strang_core(ws, dt_step)
            full = ws.psi.copy()
            ws.psi = saved.copy()
            strang_core(ws, dt_step / 2)
            strang_core(ws, dt_step / 2)
            err = l2_change(full, ws.psi)
            if may_reject and err > tol:
                ws.psi = saved
                dt = max(dt_step * max(0.5, 0.9 * (tol / err) ** (1 / 3)), dt_min)
                n_rej += 1
                continue
        else:
            old = ws.psi.copy() if not clamped else None
            prev = (fsal, fsal_dt)
            if fsal and abs(fsal_dt - dt_step) < 1e-14:
                hp(dt_step)
            elif fsal:
                hp(fsal_dt / 2)
                hp(dt_step / 2)
            else:
                hp(dt_step / 2)
            fsal = False
            ws.psi = kinetic_step(ws.psi, kinetic_phase(ws.grid, dt_step, False))
            err = 0.0
            if may_reject:
                err = l2_change(ws.psi, old)
                if err > tol:
                    ws.psi = old
                    fsal, fsal_dt = prev
                    dt = max(dt_step * max(0.5, 0.9 * math.sqrt(tol / err)), dt_min)
                    n_rej += 1
                    continue
            fsal, fsal_dt = True, dt_step
        apply_loss(ws, dt_step)
        ws.t += dt_step
        ws.step += 1
        n_acc += 1
        if not clamped:
            if error_mode == "richardson":
                fac = min(2.0, 0.9 * (tol / err) ** (1 / 3)) if err > 1e-300 else 2.0
            else:
                if not may_reject:
                    err = l2_change(ws.psi, old)
                fac = min(2.0, 0.9 * math.sqrt(tol / err)) if err > 1e-300 else 2.0
            dt = min(max(dt * fac, dt_min), dt_max)
        if ws.t >= next_save - 1e-14:
            if fsal:
                hp(fsal_dt / 2)
                fsal = False
            _snapshot(ws, rec)
            next_save += save_interval
    if fsal:
        hp(fsal_dt / 2)
    if not rec["times"] or abs(rec["times"][-1] - ws.t) > 1e-12:
        _snapshot(ws, rec)
    return dict(result=rec, n_accepted=n_acc, n_rejected=n_rej, final_dt=dt)


def relative_l2(a, b):
    """Per-component relative L2 error max_c ||a_c - b_c|| / ||b_c|| (components
    with ||b_c|| == 0 use the absolute error scaled by the total norm)."""
    tot = math.sqrt(float(np.sum(np.abs(b) ** 2))) or 1.0
    errs = []
    for c in range(b.shape[0]):
        nb = math.sqrt(float(np.sum(np.abs(b[c]) ** 2)))
        nd = math.sqrt(float(np.sum(np.abs(a[c] - b[c]) ** 2)))
        errs.append(nd / (nb if nb > 1e-12 * tot else tot))
    return max(errs)
