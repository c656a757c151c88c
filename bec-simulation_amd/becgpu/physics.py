"""Host-side model setup mirroring SpinorBEC.jl's L1 layer (setup only; no
time stepping happens here).

grid.jl, types.jl, spin_matrices.jl, clebsch_gordan.jl, interactions.jl,
zeeman.jl, potentials.jl, initialization.jl, thomas_fermi.jl, atoms.jl,
units.jl, diagnostics.jl (classify_phase).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

# ---------------------------------------------------------------------------
# Units / atoms (units.jl:3-7, atoms.jl)
# ---------------------------------------------------------------------------
HBAR = 1.054571817e-34
AMU = 1.66053906660e-27
BOHR_RADIUS = 5.29177210903e-11
BOHR_MAGNETON = 9.2740100783e-24


@dataclass(frozen=True)
class AtomSpecies:
    """types.jl:276-320 (only the fields the hot path reads)."""
    name: str
    mass: float
    F: int
    a0: float = 0.0
    a2: float = 0.0
    mu_mag: float = 0.0
    g_F: float = 0.0

    @property
    def a_s(self):
        return (self.a0 + 2 * self.a2) / 3 if self.F == 1 else self.a0


Rb87 = AtomSpecies("87Rb", 86.909180527 * AMU, 1, 101.8 * BOHR_RADIUS, 100.4 * BOHR_RADIUS, 0.0, -0.5)
Na23 = AtomSpecies("23Na", 22.9897692820 * AMU, 1, 50.0 * BOHR_RADIUS, 55.0 * BOHR_RADIUS, 0.0, -0.5)
Eu151 = AtomSpecies("151Eu", 150.919857 * AMU, 6, 110.0 * BOHR_RADIUS, 0.0,
                    1.9934 * 3.5 * BOHR_MAGNETON, 1.9934 * 7.0 / 12.0)


# ---------------------------------------------------------------------------
# Grid (grid.jl:1-57, types.jl:1-23)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class GridConfig:
    n_points: tuple
    box_size: tuple

    def __post_init__(self):
        n = tuple(int(v) for v in np.atleast_1d(self.n_points))
        L = tuple(float(v) for v in np.atleast_1d(self.box_size))
        if len(n) != len(L) or not 1 <= len(n) <= 3:
            raise ValueError("n_points and box_size must have the same length 1..3")
        if not all(v > 0 and v % 2 == 0 for v in n):
            raise ValueError("n_points must be positive even integers")
        if not all(v > 0 for v in L):
            raise ValueError("box_size must be positive")
        object.__setattr__(self, "n_points", n)
        object.__setattr__(self, "box_size", L)

    @property
    def ndim(self):
        return len(self.n_points)


@dataclass(frozen=True, eq=False)
class Grid:
    config: GridConfig
    x: tuple
    dx: tuple
    k: tuple
    dk: tuple

    @property
    def ndim(self):
        return self.config.ndim

    @property
    def shape(self):
        """numpy (C-order) shape of one component: (nz, ny, nx)."""
        return tuple(reversed(self.config.n_points))

    @property
    def k_squared(self):
        k2 = np.zeros(self.shape)
        nd = self.ndim
        for d in range(nd):
            sh = [1] * nd
            sh[nd - 1 - d] = self.config.n_points[d]
            k2 = k2 + (self.k[d] ** 2).reshape(sh)
        return k2


def _fftfreq(n, fs):
    """Julia FFTW.fftfreq(n, fs): j * (fs/n), j = 0..n/2-1, -n/2..-1."""
    mult = fs / n
    j = np.concatenate([np.arange(0, n // 2), np.arange(-(n // 2), 0)]) if n % 2 == 0 else \
        np.concatenate([np.arange(0, (n + 1) // 2), np.arange(-(n // 2), 0)])
    return j * mult


def make_grid(config: GridConfig) -> Grid:
    """grid.jl:1-23: cell-centred x; k = fftfreq(n, n*2pi/L)."""
    xs, dxs, ks, dks = [], [], [], []
    for n, L in zip(config.n_points, config.box_size):
        dx = L / n
        xs.append(np.linspace(-L / 2 + dx / 2, L / 2 - dx / 2, n))
        dxs.append(dx)
        dk = 2 * math.pi / L
        ks.append(_fftfreq(n, n * dk))
        dks.append(dk)
    return Grid(config, tuple(xs), tuple(dxs), tuple(ks), tuple(dks))


def cell_volume(grid: Grid) -> float:
    return float(np.prod(grid.dx))


# ---------------------------------------------------------------------------
# Spin system and matrices (types.jl:227-252, spin_matrices.jl:1-50)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class SpinSystem:
    F: int

    def __post_init__(self):
        if self.F < 0:
            raise ValueError("F must be non-negative")

    @property
    def n_components(self):
        return 2 * self.F + 1

    @property
    def m_values(self):
        return list(range(self.F, -self.F - 1, -1))


@lru_cache(maxsize=None)
def spin_matrices(F: int):
    D = 2 * F + 1
    m = [F - c for c in range(D)]
    Fp = np.zeros((D, D), complex)
    for j in range(D):
        for i in range(D):
            if m[i] == m[j] + 1:
                Fp[i, j] = math.sqrt(F * (F + 1) - m[j] * (m[j] + 1))
    Fm = Fp.conj().T
    Fx = (Fp + Fm) / 2
    Fy = (Fp - Fm) / 2j
    Fz = np.diag(m).astype(complex)
    vals, vecs = np.linalg.eigh(Fy)          # LAPACK, ascending (spin_matrices.jl:38)
    if not np.all(np.abs(np.diff(vals) - 1.0) < 1e-10):
        raise AssertionError(f"Fy eigenvalues not uniformly spaced: {vals}")
    return dict(F=F, D=D, Fx=Fx, Fy=Fy, Fz=Fz, Fp=Fp, Fm=Fm, fy_vals=vals, fy_vecs=vecs)


# ---------------------------------------------------------------------------
# Angular momentum algebra (clebsch_gordan.jl:1-180), Racah formulas in log space
# ---------------------------------------------------------------------------
def _log_factorial(n: int) -> float:
    if n < 0:
        return -math.inf
    return math.lgamma(n + 1)


def _log_triangle(a, b, c):
    ab, ac, bc = a + b - c, a - b + c, -a + b + c
    if ab < 0 or ac < 0 or bc < 0:
        return -math.inf
    return _log_factorial(ab) + _log_factorial(ac) + _log_factorial(bc) - _log_factorial(a + b + c + 1)


def wigner_3j(j1, j2, j3, m1, m2, m3) -> float:
    if m1 + m2 + m3 != 0:
        return 0.0
    lt = _log_triangle(j1, j2, j3)
    if lt == -math.inf:
        return 0.0
    ln = sum(_log_factorial(v) for v in (j1 + m1, j1 - m1, j2 + m2, j2 - m2, j3 + m3, j3 - m3))
    tmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    tmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    if tmin > tmax:
        return 0.0
    s = 0.0
    for t in range(tmin, tmax + 1):
        ld = (_log_factorial(t) + _log_factorial(j1 + j2 - j3 - t) + _log_factorial(j1 - m1 - t)
              + _log_factorial(j2 + m2 - t) + _log_factorial(j3 - j2 + m1 + t)
              + _log_factorial(j3 - j1 - m2 + t))
        s += (-1.0 if t % 2 else 1.0) * math.exp(-ld)
    phase = 1.0 if (j1 - j2 - m3) % 2 == 0 else -1.0
    return phase * math.exp(0.5 * (lt + ln)) * s


def clebsch_gordan(j1, m1, j2, m2, J, M) -> float:
    if m1 + m2 != M:
        return 0.0
    phase = 1.0 if (j1 - j2 + M) % 2 == 0 else -1.0
    return phase * math.sqrt(2 * J + 1) * wigner_3j(j1, j2, J, m1, m2, -M)


def wigner_6j(j1, j2, j3, j4, j5, j6) -> float:
    t1, t2 = _log_triangle(j1, j2, j3), _log_triangle(j1, j5, j6)
    t3, t4 = _log_triangle(j4, j2, j6), _log_triangle(j4, j5, j3)
    if -math.inf in (t1, t2, t3, t4):
        return 0.0
    lts = 0.5 * (t1 + t2 + t3 + t4)
    tmin = max(j1 + j2 + j3, j1 + j5 + j6, j4 + j2 + j6, j4 + j5 + j3)
    tmax = min(j1 + j2 + j4 + j5, j2 + j3 + j5 + j6, j1 + j3 + j4 + j6)
    if tmin > tmax:
        return 0.0
    s = 0.0
    for t in range(tmin, tmax + 1):
        ld = (_log_factorial(t - j1 - j2 - j3) + _log_factorial(t - j1 - j5 - j6)
              + _log_factorial(t - j4 - j2 - j6) + _log_factorial(t - j4 - j5 - j3)
              + _log_factorial(j1 + j2 + j4 + j5 - t) + _log_factorial(j2 + j3 + j5 + j6 - t)
              + _log_factorial(j1 + j3 + j4 + j6 - t))
        s += (-1.0 if t % 2 else 1.0) * math.exp(_log_factorial(t + 1) - ld)
    return math.exp(lts) * s


@lru_cache(maxsize=None)
def precompute_cg_table(F: int):
    """clebsch_gordan.jl:166-180: {(l, M, m1, m2): <F m1; F m2 | l M>}, even l."""
    t = {}
    for l in range(0, 2 * F + 1, 2):
        for M in range(-l, l + 1):
            for m1 in range(-F, F + 1):
                m2 = M - m1
                if abs(m2) > F:
                    continue
                v = clebsch_gordan(F, m1, F, m2, l, M)
                if abs(v) >= 1e-15:
                    t[(l, M, m1, m2)] = v
    return t


@dataclass(frozen=True)
class RamanCoupling:
    """types.jl:179-183: two-photon Raman coupling H_R = delta Fz + Omega_R/2 (e^{i k.r} F+ +
    e^{-i k.r} F-) (raman.jl:1-6); k_eff has one entry per grid axis."""
    Omega_R: float
    delta: float
    k_eff: tuple


@dataclass(frozen=True)
class LossParams:
    """types.jl:291-296: dipolar relaxation rate and three-body loss."""
    gamma_dr: float
    L3: float = 0.0


def dipolar_relaxation_rates(F: int, gamma_dr: float):
    """losses.jl:60-82: gamma_m = gamma_dr * sum_{q=-1,-2} CG(F,m;2,q|F,m+q)^2 / Z, Z the
    mean raw sum (so the mean rate is gamma_dr; m = -F is stable)."""
    D = 2 * F + 1
    raw = []
    for c in range(D):
        m = F - c
        acc = 0.0
        for q in (-1, -2):
            mp = m + q
            if abs(mp) > F:
                continue
            v = clebsch_gordan(F, m, 2, q, F, mp)
            acc += v * v
        raw.append(acc)
    tot = sum(raw)
    if tot < 1e-30:
        return [0.0] * D
    Z = tot / D
    return [gamma_dr * r / Z for r in raw]


def cn_to_gS(F, c_dict):
    """clebsch_gordan.jl:117-141."""
    ev = list(range(0, 2 * F + 1, 2))
    return {S: sum((2 * k + 1) * wigner_6j(F, F, k, F, F, S) * c_dict.get(k, 0.0) for k in ev)
            for S in ev}


def c0c1_to_gS(F, c0, c1):
    """interactions.jl:123-128."""
    return {S: c0 + c1 * (S * (S + 1) - 2 * F * (F + 1)) / 2 for S in range(0, 2 * F + 1, 2)}


def c_extra_to_delta_gS(F, c_extra):
    """interactions.jl:140-154 (c_extra[i] = c_{i+1}, 1-based i)."""
    c = {}
    for idx, val in enumerate(c_extra, start=1):
        k = idx + 1
        if abs(val) > 1e-30 and k % 2 == 0 and k <= 2 * F:
            c[k] = val
    return cn_to_gS(F, c) if c else {}


# ---------------------------------------------------------------------------
# Parameters (types.jl)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class InteractionParams:
    c0: float
    c1: float
    c_lhy: float = 0.0
    c_extra: tuple = ()

    def get_cn(self, n):
        if n == 0:
            return self.c0
        if n == 1:
            return self.c1
        idx = n - 1
        return self.c_extra[idx - 1] if idx <= len(self.c_extra) else 0.0


@dataclass(frozen=True)
class ZeemanParams:
    p: float = 0.0
    q: float = 0.0


def zeeman_energies(z: ZeemanParams, F: int):
    """zeeman.jl:15-19."""
    return [(-z.p * m + z.q * m * m) for m in range(F, -F - 1, -1)]


@dataclass(frozen=True)
class HarmonicTrap:
    omega: tuple

    def __init__(self, *omega):
        if len(omega) == 1 and np.ndim(omega[0]) == 1:
            omega = tuple(omega[0])
        object.__setattr__(self, "omega", tuple(float(w) for w in omega))


@dataclass(frozen=True)
class NoPotential:
    pass


@dataclass(frozen=True)
class GravityPotential:
    """types.jl:199-209: V = g x_axis, axis 1-based (potentials.jl:17-24)."""
    g: float
    axis: int


@dataclass(frozen=True)
class CompositePotential:
    """types.jl:211-213: the sum of its components (potentials.jl:26-32)."""
    components: tuple


def evaluate_potential(pot, grid: Grid):
    """potentials.jl:1-32 (harmonic, gravity, composite; numpy C order (z, y, x))."""
    V = np.zeros(grid.shape)
    if isinstance(pot, NoPotential) or pot is None:
        return V
    nd = grid.ndim

    def axis_field(d, vals):
        sh = [1] * nd
        sh[nd - 1 - d] = grid.config.n_points[d]
        return np.asarray(vals).reshape(sh)

    if isinstance(pot, GravityPotential):
        if not 1 <= pot.axis <= nd:
            raise ValueError(f"axis must be between 1 and {nd}")
        return V + axis_field(pot.axis - 1, pot.g * np.asarray(grid.x[pot.axis - 1]))
    if isinstance(pot, CompositePotential):
        for c in pot.components:
            V = V + evaluate_potential(c, grid)
        return V
    if not isinstance(pot, HarmonicTrap):
        raise NotImplementedError(f"potential {type(pot).__name__} is not built")
    s = np.zeros(grid.shape)
    for d in range(nd):
        s = s + axis_field(d, pot.omega[d] ** 2 * grid.x[d] ** 2)
    return 0.5 * s


@dataclass(frozen=True)
class TimeDependentZeeman:
    """types.jl:173-175: B_func(t) -> ZeemanParams, evaluated at the state's time by every
    half potential step and energy (zeeman.jl:20-21)."""
    B_func: object


def zeeman_at(z, t: float):
    """zeeman.jl:20-21."""
    return z.B_func(t) if isinstance(z, TimeDependentZeeman) else z


def interaction_params_from_constraint(*, c_total: float, c1_ratio: float = 0.0, F: int,
                                       c_extra=()):
    """interactions.jl:182-187: c0 = c_total / (1 + F^2 c1_ratio), c1 = c1_ratio c0."""
    c0 = c_total / (1.0 + F ** 2 * c1_ratio)
    return InteractionParams(c0, c1_ratio * c0, 0.0, tuple(float(v) for v in c_extra))


@dataclass(frozen=True)
class SimParams:
    """types.jl:408-437."""
    dt: float
    n_steps: int
    imaginary_time: bool = False
    normalize_every: int = None
    save_every: int = None

    def __post_init__(self):
        if not self.dt > 0:
            raise ValueError("dt must be positive")
        if not self.n_steps > 0:
            raise ValueError("n_steps must be positive")
        if self.normalize_every is None:
            object.__setattr__(self, "normalize_every", 1 if self.imaginary_time else 0)
        if self.save_every is None:
            object.__setattr__(self, "save_every", max(1, self.n_steps // 100))


@dataclass(frozen=True)
class AdaptiveDtParams:
    """types.jl:343-361."""
    dt_init: float = 0.001
    dt_min: float = 1e-5
    dt_max: float = 0.01
    tol: float = 1e-3
    error_mode: str = "step_change"

    def __post_init__(self):
        if not self.dt_init > 0:
            raise ValueError("dt_init must be positive")
        if not self.dt_min > 0:
            raise ValueError("dt_min must be positive")
        if not self.dt_max >= self.dt_min:
            raise ValueError("dt_max must be >= dt_min")
        if not self.tol > 0:
            raise ValueError("tol must be positive")
        if self.error_mode not in ("step_change", "richardson"):
            raise ValueError(f"error_mode must be step_change or richardson, got {self.error_mode}")


# ---------------------------------------------------------------------------
# Initial states (initialization.jl:1-79, thomas_fermi.jl:15-77)
# ---------------------------------------------------------------------------
class SplitMix64:
    """Seeded generator for the random initial states and noise.  (The
    reference's Julia MersenneTwister stream cannot be reproduced outside
    Julia; every consumer of a random psi receives the same array.)"""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def next_u64_array(self, n):
        """The next n outputs at once (state k*gamma ahead, then the same mixing)."""
        g = np.uint64(0x9E3779B97F4A7C15)
        with np.errstate(over="ignore"):
            z = np.uint64(self.s) + g * np.arange(1, n + 1, dtype=np.uint64)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.s = (self.s + n * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        return z

    def uniform(self, n):
        """U(0,1): the top 53 bits of each output, centred in its interval."""
        return ((self.next_u64_array(n) >> np.uint64(11)).astype(np.float64) + 0.5) * \
            (1.0 / 9007199254740992.0)

    def randn_complex(self, n):
        """Complex normal with E|z|^2 = 1 (Box-Muller; re, im ~ N(0, 1/2))."""
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        r = np.sqrt(-np.log(u1))
        return r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2)


def _gaussian(grid: Grid, sigma):
    s = np.zeros(grid.shape)
    nd = grid.ndim
    for d in range(nd):
        sh = [1] * nd
        sh[nd - 1 - d] = grid.config.n_points[d]
        s = s + (grid.x[d] ** 2 / (2 * sigma[d] ** 2)).reshape(sh)
    return np.exp(-s)


def init_psi(grid: Grid, sys: SpinSystem, state="polar", seed=42):
    """initialization.jl:1-67 -> numpy array (D, nz, ny, nx) complex128."""
    D, F = sys.n_components, sys.F
    psi = np.zeros((D,) + grid.shape, complex)
    g = _gaussian(grid, [L / 8 for L in grid.config.box_size])
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
    elif state == "random":
        z = SplitMix64(seed).randn_complex(psi.size).reshape(psi.shape)
        psi = z * g[None]
    else:
        raise ValueError(f"Unknown initial state: {state}")
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * cell_volume(grid))
    return psi


def init_psi_thomas_fermi(grid: Grid, sys: SpinSystem, potential, c0, N_target=1.0):
    """thomas_fermi.jl:15-77 (polar component, bisection tol 1e-12)."""
    V = evaluate_potential(potential, grid)
    dV = cell_volume(grid)

    def tf_norm(mu):
        val = mu - V
        return float(np.sum(np.where(val > 0, val / c0, 0.0))) * dV

    Vmin, Vmax = float(V.min()), float(V.max())
    lo, hi = Vmin, Vmin + (Vmax - Vmin) * 2.0
    while tf_norm(hi) < N_target and hi < Vmax * 100:
        hi *= 2.0
    mu = None
    for _ in range(200):
        mid = (lo + hi) / 2
        Nm = tf_norm(mid)
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
    D = sys.n_components
    psi = np.zeros((D,) + grid.shape, complex)
    psi[(D + 1) // 2 - 1] = np.sqrt(n)
    nrm = math.sqrt(np.sum(np.abs(psi) ** 2) * dV)
    if nrm > 0:
        psi = psi / nrm * math.sqrt(N_target)
    return psi


def add_noise(psi, amplitude, seed, components=None, grid=None):
    """Seeded complex Gaussian noise of the given amplitude in the listed
    components (the non-dominant ones by default), renormalised to the
    incoming norm (experiment_runner.jl:157-176 semantics, own generator)."""
    D = psi.shape[0]
    if components is None:
        pops = np.sum(np.abs(psi.reshape(D, -1)) ** 2, axis=1)
        dom = int(np.argmax(pops))
        components = [c for c in range(D) if c != dom]
    out = psi.copy()
    gen = SplitMix64(seed)
    n0 = np.sum(np.abs(psi) ** 2)
    for c in components:
        out[c] = out[c] + amplitude * gen.randn_complex(psi[c].size).reshape(psi[c].shape)
    out *= math.sqrt(n0 / np.sum(np.abs(out) ** 2))
    return out


# ---------------------------------------------------------------------------
# classify_phase (diagnostics.jl:166-212) -- host analysis used by scan_continuation
# ---------------------------------------------------------------------------
def spin_density_vector(psi, F):
    """observables.jl:61-88."""
    D = 2 * F + 1
    m = np.array([F - c for c in range(D)], float)
    fz = np.tensordot(m, np.abs(psi) ** 2, axes=(0, 0))
    fxy = np.zeros(psi.shape[1:], complex)
    for c in range(1, D):
        fxy += math.sqrt(F * (F + 1) - m[c] * (m[c] + 1)) * np.conj(psi[c - 1]) * psi[c]
    return fxy.real, fxy.imag, fz


def classify_phase(psi, F, grid: Grid):
    D = 2 * F + 1
    dV = cell_volume(grid)
    n = np.sum(np.abs(psi) ** 2, axis=0)
    n_sum = float(np.sum(n)) * dV
    n_sq = float(np.sum(n ** 2)) * dV
    if n_sum < 1e-30:
        return dict(spin_order=0.0, nematic_order=0.0, channel_weights={}, phase="vacuum",
                    magnetization_density=0.0)
    fx, fy, fz = spin_density_vector(psi, F)
    spin_order = float(np.sum(fx ** 2 + fy ** 2 + fz ** 2)) * dV / (float(F) ** 2 * n_sq)
    cg = precompute_cg_table(F)
    weights = {}
    for S in range(0, 2 * F + 1, 2):
        w = 0.0
        for M in range(-S, S + 1):
            A = np.zeros(psi.shape[1:], complex)
            for m1 in range(-F, F + 1):
                m2 = M - m1
                if abs(m2) > F:
                    continue
                v = cg.get((S, M, m1, m2), 0.0)
                if abs(v) < 1e-15:
                    continue
                A += v * psi[F - m1] * psi[F - m2]
            w += float(np.sum(np.abs(A) ** 2)) * dV
        weights[S] = w
    tot = sum(weights.values())
    cw = {S: (w / tot if tot > 0 else 0.0) for S, w in weights.items()}
    nematic_order = weights.get(0, 0.0) / (n_sq / D)
    Mz = sum((F - c) * float(np.sum(np.abs(psi[c]) ** 2)) * dV for c in range(D)) / n_sum
    if spin_order > 0.9:
        phase = "ferromagnetic"
    elif nematic_order > 0.9:
        phase = "polar" if F == 1 else "nematic"
    elif cw.get(2 * F, 0.0) > 0.5:
        phase = "cyclic"
    else:
        phase = "mixed"
    return dict(spin_order=spin_order, nematic_order=nematic_order, channel_weights=cw,
                phase=phase, magnetization_density=Mz)
