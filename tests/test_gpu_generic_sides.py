"""Grid sides without a built radix plan (the runtime-length line passes) against the FP64
oracle.

The reference takes any positive even side (types.jl:12) and plans FFTW for it
(grid.jl:37-53; zero-padded DDI on twice the side, ddi_padded.jl:6-49).  The sides here
have factors 2, 3, 5, 7 and larger odd primes (11, 13, 101), alone and beside built sides
on the other axes; zero-padded DDI lines of twice such sides, and of 224 and 512 points,
whose doubles have no built plan; 13 components on 512-point x rows (more than one block
of built-plan threads holds); FP32 bulk; the spectral gradient (every mode, Nyquist
included) to 1e-12.  Sides with a large prime factor (214 = 2 107, 218, 226, 254, 502 and
the padded 452) run Bluestein's algorithm instead of a radix-p pass, on x lines, y/z lines
and DDI spectra lines.  Bars: per-component relative L2 1e-10 and total energy 1e-12 in FP64.
"""
import math

import numpy as np
import pytest

from oracle import spinor_oracle as O
import becgpu as G
from becgpu import api as A
from helpers import grids, pair, random_psi, rel_l2

pytestmark = pytest.mark.gpu

PSI_TOL = 1e-10
E_TOL = 1e-12


def _psi(n, L, F, seed):
    og, _ = grids(n, L)
    D = 2 * F + 1
    psi = O.init_psi(og, F, "uniform") * O.gaussian(og, tuple(0.3 * l for l in L))[None]
    psi = psi + 0.05 * random_psi(og.shape, D, seed)
    return psi / math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)


def _cextra(seed):
    r = np.random.default_rng(seed)
    ce = [0.0] * 11
    for k in range(2, 13, 2):
        ce[k - 2] = float(r.uniform(-1, 1))
    return ce


def _check(ows, gws, nsteps, psi_tol=PSI_TOL, e_tol=E_TOL):
    for _ in range(nsteps):
        O.split_step(ows)
    A.split_step(gws, nsteps)
    got = gws.get_psi()
    err = max(rel_l2(got[c], ows.psi[c]) for c in range(got.shape[0]))
    eo = O.total_energy(ows)
    de = abs(A.total_energy(gws) - eo) / abs(eo)
    print("psi rel L2", err, "energy rel", de)
    assert err <= psi_tol
    assert de <= e_tol


@pytest.mark.parametrize("n", [6, 10, 22, 26, 72, 100, 120, 144, 200, 202])
def test_1d_lines(n):
    L = (16.0,)
    psi = _psi((n,), L, 1, n)
    for itp in (False, True):
        ows, gws = pair((n,), L, 1, psi, 0.004, itp, omega=(1.0,), p=0.1, q=0.2, c0=8.0,
                        c1=-0.4)
        _check(ows, gws, 20)


@pytest.mark.parametrize("n,L,F,itp", [((72, 100), (12.0, 16.0), 1, True),
                                       ((120, 26), (16.0, 8.0), 2, False),
                                       ((144, 64), (20.0, 12.0), 1, False),
                                       ((64, 144), (12.0, 20.0), 2, True)])
def test_2d(n, L, F, itp):
    psi = _psi(n, L, F, 5)
    ows, gws = pair(n, L, F, psi, 0.004, itp, omega=(1.0, 1.3), p=0.1, q=0.2, c0=8.0, c1=-0.4,
                    c_lhy=0.2)
    _check(ows, gws, 15)


@pytest.mark.parametrize("n,L,F,itp", [((12, 10, 18), (8.0, 7.0, 10.0), 1, False),
                                       ((20, 14, 8), (10.0, 8.0, 6.0), 2, True),
                                       ((10, 12, 6), (7.0, 8.0, 5.0), 6, False),
                                       ((30, 22, 16), (10.0, 9.0, 8.0), 3, True)])
def test_3d(n, L, F, itp):
    psi = _psi(n, L, F, 7)
    kw = dict(c_extra=_cextra(3)) if F == 6 else dict(c1=-0.4)
    ows, gws = pair(n, L, F, psi, 0.002, itp, omega=(1.0, 1.1, 1.2), q=0.1, c0=8.0, **kw)
    _check(ows, gws, 10)


@pytest.mark.parametrize("padding", [False, True])
@pytest.mark.parametrize("secular", [False, True])
@pytest.mark.parametrize("n,L", [((100,), (20.0,)), ((72, 30), (14.0, 10.0)),
                                 ((18, 12, 10), (9.0, 8.0, 7.0)), ((16, 18, 8), (8.0, 9.0, 6.0))])
def test_ddi(n, L, secular, padding):
    """Half spectra of 51, 37, 10 bins; padded lines of 200, 144, 60, 36, 24, 20 points."""
    psi = _psi(n, L, 1, 11)
    ows, gws = pair(n, L, 1, psi, 0.002, False, omega=(1.0,) * len(n), c0=10.0, c1=-0.5,
                    ddi=True, cdd=2.0, secular=secular, padding=padding)
    _check(ows, gws, 8)


@pytest.mark.parametrize("n,L", [((224, 8), (20.0, 6.0)), ((512, 4), (30.0, 4.0)),
                                 ((56, 224), (9.0, 20.0))])
def test_ddi_padded_without_built_double(n, L):
    """Zero-padded DDI where twice the side (448, 1024) has no built plan: the x pass (or the
    y pass) runs the runtime-length kernel, the other axis keeps its built lines."""
    psi = _psi(n, L, 1, 13)
    ows, gws = pair(n, L, 1, psi, 0.002, True, omega=(1.0,) * len(n), c0=10.0, c1=-0.5,
                    ddi=True, cdd=2.0, padding=True)
    _check(ows, gws, 6)


def test_spin6_tensor_ddi_512_rows():
    """13 components on 512-point x rows: more threads than one built-plan block holds, so
    the x pass runs the runtime-length kernel (one row per block)."""
    n, L = (512,), (30.0,)
    psi = _psi(n, L, 6, 17)
    ows, gws = pair(n, L, 6, psi, 1e-3, False, omega=(1.0,), c0=10.0, c_extra=_cextra(6),
                    ddi=True, cdd=1.0)
    _check(ows, gws, 3)


@pytest.mark.parametrize("padding", [False, True])
def test_spin6_tensor_ddi_2d(padding):
    n, L = (30, 18), (10.0, 8.0)
    psi = _psi(n, L, 6, 19)
    ows, gws = pair(n, L, 6, psi, 1e-3, True, omega=(1.0, 1.0), c0=10.0, c_extra=_cextra(7),
                    ddi=True, cdd=1.0, padding=padding)
    _check(ows, gws, 4)


@pytest.mark.parametrize("n,L", [((100,), (10.0,)), ((72, 30), (12.0, 8.0)),
                                 ((18, 14, 22), (9.0, 14.0, 11.0)), ((6, 202), (4.0, 20.0))])
def test_fft_gradient(n, L):
    og, gg = grids(n, L)
    f = np.random.default_rng(5).normal(size=og.shape)
    ref = O.fft_gradient(f, og)
    got = G.fft_gradient(f, gg)
    for d in range(og.ndim):
        err = np.max(np.abs(got[d] - ref[d])) / np.max(np.abs(ref[d]))
        print(n, d, err)
        assert err < 1e-12


def test_yoshida_spin2_ddi():
    n, L = (30, 18, 12), (10.0, 8.0, 7.0)
    psi = _psi(n, L, 2, 15)
    ows, gws = pair(n, L, 2, psi, 1e-3, False, omega=(1.0, 1.0, 1.0), c0=8.0, c1=-0.7,
                    ddi=True, cdd=1.0)
    for _ in range(2):
        O.yoshida_core(ows, 1e-3)
        A.yoshida_core(gws, 1e-3)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    eo = O.total_energy(ows)
    assert abs(A.total_energy(gws) - eo) <= E_TOL * abs(eo)


def test_fp32_3d():
    """FP32 bulk on 72 x 50 x 20 with DDI: error stated against the FP64 oracle."""
    n, L = (72, 50, 20), (14.0, 12.0, 8.0)
    psi = _psi(n, L, 1, 17)
    ows, gws = pair(n, L, 1, psi, 0.002, False, omega=(1.0, 1.0, 1.0), q=0.1, c0=10.0, c1=0.5,
                    ddi=True, cdd=1.0, precision="fp32")
    _check(ows, gws, 10, psi_tol=1e-5, e_tol=1e-6)


def test_ground_state_cfg_side_100():
    """find_ground_state on a 100-point line (ITP with the check interval, normalisation
    and convergence test) against the oracle's driver."""
    n, L = (100,), (20.0,)
    og, gg = grids(n, L)
    psi = O.init_psi(og, 1, "polar")
    r = A.find_ground_state(grid=gg, atom=G.Rb87, interactions=G.InteractionParams(10.0, -0.5),
                            potential=G.HarmonicTrap(1.0), dt=0.005, n_steps=400, tol=1e-10,
                            initial_state="polar")
    ows = O.make_workspace(og, 1, psi, 0.005, True, omega=(1.0,), c0=10.0, c1=-0.5, n_steps=400)
    o = O.find_ground_state(ows, 400, 1e-10)
    assert r["converged"] == o["converged"] and r["steps"] == o["steps"]
    assert abs(r["energy"] - o["energy"]) <= 1e-12 * abs(o["energy"])
    assert rel_l2(r["workspace"].get_psi(), ows.psi) <= 1e-10



def _bluestein(*sides):
    """The sides run Bluestein's algorithm (bec_line_plan 2)."""
    from becgpu import _lib as L
    for n in sides:
        assert L.lib().bec_line_plan(n) == 2, n


@pytest.mark.parametrize("n", [214, 226, 254, 502])
def test_bluestein_1d_lines(n):
    _bluestein(n)
    L = (16.0,)
    psi = _psi((n,), L, 1, n)
    for itp in (False, True):
        ows, gws = pair((n,), L, 1, psi, 0.004, itp, omega=(1.0,), p=0.1, q=0.2, c0=8.0,
                        c1=-0.4)
        _check(ows, gws, 20)


@pytest.mark.parametrize("n,L,F,itp", [((254, 30), (20.0, 8.0), 1, True),
                                       ((30, 226), (8.0, 20.0), 2, False),
                                       ((218, 12), (20.0, 6.0), 1, False)])
def test_bluestein_2d(n, L, F, itp):
    _bluestein(max(n))
    psi = _psi(n, L, F, 21)
    ows, gws = pair(n, L, F, psi, 0.004, itp, omega=(1.0, 1.3), p=0.1, q=0.2, c0=8.0, c1=-0.4)
    _check(ows, gws, 15)


@pytest.mark.parametrize("n,L,F,itp", [((10, 8, 254), (6.0, 5.0, 20.0), 1, True),
                                       ((12, 226, 6), (7.0, 20.0, 5.0), 2, False)])
def test_bluestein_3d(n, L, F, itp):
    _bluestein(max(n))
    psi = _psi(n, L, F, 23)
    ows, gws = pair(n, L, F, psi, 0.002, itp, omega=(1.0, 1.1, 1.2), q=0.1, c0=8.0, c1=-0.4)
    _check(ows, gws, 10)


@pytest.mark.parametrize("padding", [False, True])
@pytest.mark.parametrize("n,L", [((226,), (20.0,)), ((8, 226), (6.0, 20.0))])
def test_bluestein_ddi(n, L, padding):
    """Spectra of 226-point lines; zero-padded lines of 452 points (Bluestein) on the x axis
    (the x pass's auxiliary lines) and on the y axis (the DDI spectra pass)."""
    _bluestein(226, 452)
    psi = _psi(n, L, 1, 25)
    ows, gws = pair(n, L, 1, psi, 0.002, False, omega=(1.0,) * len(n), c0=10.0, c1=-0.5,
                    ddi=True, cdd=2.0, padding=padding)
    _check(ows, gws, 8)


def test_bluestein_fp32():
    """FP32 bulk on 254 x 30 with DDI: error stated against the FP64 oracle (the Bluestein
    work lines are FP64 in both precisions)."""
    n, L = (254, 30), (20.0, 8.0)
    psi = _psi(n, L, 1, 27)
    ows, gws = pair(n, L, 1, psi, 0.002, False, omega=(1.0, 1.0), q=0.1, c0=10.0, c1=0.5,
                    ddi=True, cdd=1.0, precision="fp32")
    _check(ows, gws, 10, psi_tol=1e-5, e_tol=1e-6)


@pytest.mark.parametrize("n,L", [((254,), (20.0,)), ((30, 226), (8.0, 20.0)),
                                 ((6, 8, 218), (5.0, 6.0, 20.0))])
def test_bluestein_gradient(n, L):
    _bluestein(max(n))
    og, gg = grids(n, L)
    f = np.random.default_rng(9).normal(size=og.shape)
    ref = O.fft_gradient(f, og)
    got = G.fft_gradient(f, gg)
    for d in range(og.ndim):
        err = np.max(np.abs(got[d] - ref[d])) / np.max(np.abs(ref[d]))
        print(n, d, err)
        assert err < 1e-12
