"""Grid sides of 3 * 2^k, 5 * 2^k and 7 * 2^k points (radix-3, -5 and -7 Stockham passes)
against the FP64 oracle.

The reference takes any positive even side (grid.jl: GridConfig validation, FFTW plans of
any length).  Built here: 2^k (8..512), 3 * 2^k (24..192), 5 * 2^k (40..160) and 7 * 2^k
(56..224) sides, with zero-padded DDI lines of twice the side (up to 384, 320 and 224).  Sides are mixed per axis so
that a swapped axis or a wrong radix order fails; the FFT itself is checked through the
spectral gradient (every mode, Nyquist included) to 1e-12.
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


@pytest.mark.parametrize("n", [24, 40, 48, 56, 80, 96, 112, 160, 192, 224])
def test_1d_lines(n):
    L = (16.0,)
    psi = _psi((n,), L, 1, n)
    ows, gws = pair((n,), L, 1, psi, 0.004, False, omega=(1.0,), p=0.1, q=0.2, c0=8.0, c1=-0.4)
    _check(ows, gws, 20)


@pytest.mark.parametrize("n,L,F,itp", [((48, 80), (12.0, 16.0), 2, True),
                                       ((96, 40), (16.0, 10.0), 1, False),
                                       ((160, 24), (20.0, 8.0), 1, True),
                                       ((192, 64), (20.0, 12.0), 2, False)])
def test_2d(n, L, F, itp):
    psi = _psi(n, L, F, 5)
    ows, gws = pair(n, L, F, psi, 0.004, itp, omega=(1.0, 1.3), p=0.1, q=0.2, c0=8.0, c1=-0.4,
                    c_lhy=0.2)
    _check(ows, gws, 15)


@pytest.mark.parametrize("n,L,F,itp", [((48, 40, 24), (12.0, 10.0, 8.0), 1, False),
                                       ((24, 48, 80), (8.0, 12.0, 16.0), 2, True),
                                       ((40, 24, 16), (10.0, 8.0, 6.0), 6, False),
                                       ((112, 56, 48), (16.0, 10.0, 9.0), 1, True),
                                       ((56, 16, 112), (10.0, 6.0, 16.0), 2, False)])
def test_3d(n, L, F, itp):
    psi = _psi(n, L, F, 7)
    kw = dict(c_extra=_cextra(3)) if F == 6 else dict(c1=-0.4)
    ows, gws = pair(n, L, F, psi, 0.002, itp, omega=(1.0, 1.1, 1.2), q=0.1, c0=8.0, **kw)
    _check(ows, gws, 10)


@pytest.mark.parametrize("padding", [False, True])
def test_spin6_mixed_transverse_sides(padding):
    """x side a power of two with y, z sides of 48 and 24: the fused x/site kernel (built for
    power-of-two grids) declines the shape, and the split passes take site coordinates and
    DDI spectra rows by division."""
    n, L = (32, 48, 24), (12.0, 14.0, 10.0)
    og, _ = grids(n, L)
    psi = _psi(n, L, 6, 9)
    ows, gf = pair(n, L, 6, psi, 1e-3, False, omega=(1.0, 1.0, 1.0), c0=10.0, c_extra=_cextra(4),
                   ddi=True, cdd=1.0, padding=padding)
    _, gs = pair(n, L, 6, psi, 1e-3, False, omega=(1.0, 1.0, 1.0), c0=10.0, c_extra=_cextra(4),
                 ddi=True, cdd=1.0, padding=padding)
    assert not gf.set_xsite(True)
    gs.set_xsite(False)
    gf.profile(True)
    for _ in range(2):
        O.split_step(ows)
    A.split_step(gf, 2)
    A.split_step(gs, 2)
    launches = {k: v[1] for k, v in gf.profile_read().items()}
    assert launches["xsite"] == 0 and launches["sites"] > 0, launches
    a, b = gf.get_psi(), gs.get_psi()
    assert np.array_equal(a, b), float(np.max(np.abs(a - b)))
    assert rel_l2(a, ows.psi) <= PSI_TOL
    eo = O.total_energy(ows)
    assert abs(A.total_energy(gf) - eo) <= E_TOL * abs(eo)


@pytest.mark.parametrize("padding", [False, True])
@pytest.mark.parametrize("secular", [False, True])
@pytest.mark.parametrize("n,L", [((24, 40, 48), (10.0, 12.0, 14.0)), ((80, 48), (16.0, 12.0)),
                                 ((96,), (20.0,)), ((56, 112), (10.0, 16.0))])
def test_ddi(n, L, secular, padding):
    """DDI on mixed sides: half spectra of 13, 21, 25 ... bins, padded lines of 48 ... 192."""
    psi = _psi(n, L, 1, 11)
    ows, gws = pair(n, L, 1, psi, 0.002, False, omega=(1.0,) * len(n), c0=10.0, c1=-0.5,
                    ddi=True, cdd=2.0, secular=secular, padding=padding)
    _check(ows, gws, 8)


@pytest.mark.parametrize("n,L", [((192,), (24.0,)), ((160, 32), (20.0, 8.0)), ((64, 192), (12.0, 20.0)),
                                 ((112, 56), (16.0, 10.0))])
def test_ddi_padded_long_lines(n, L):
    """Zero-padded DDI with 384-, 320- and 224-point spectral lines (x pass and y pass)."""
    psi = _psi(n, L, 1, 13)
    ows, gws = pair(n, L, 1, psi, 0.002, True, omega=(1.0,) * len(n), c0=10.0, c1=-0.5,
                    ddi=True, cdd=2.0, padding=True)
    _check(ows, gws, 6)


@pytest.mark.parametrize("n,L", [((48,), (10.0,)), ((80, 24), (12.0, 8.0)),
                                 ((40, 96, 48), (9.0, 14.0, 11.0)), ((224, 56), (20.0, 9.0))])
def test_fft_gradient(n, L):
    """Spectral gradient (fft_utils.jl:4-44): forward and inverse line FFTs of every built
    radix against numpy on a seeded random field."""
    og, gg = grids(n, L)
    f = np.random.default_rng(5).normal(size=og.shape)
    ref = O.fft_gradient(f, og)
    got = G.fft_gradient(f, gg)
    for d in range(og.ndim):
        err = np.max(np.abs(got[d] - ref[d])) / np.max(np.abs(ref[d]))
        print(n, d, err)
        assert err < 1e-12


def test_yoshida_spin2_ddi():
    n, L = (40, 24, 48), (10.0, 8.0, 12.0)
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
    """FP32 bulk on 96 x 48 x 40: error stated against the FP64 oracle."""
    n, L = (96, 48, 40), (16.0, 12.0, 10.0)
    psi = _psi(n, L, 1, 17)
    ows, gws = pair(n, L, 1, psi, 0.002, False, omega=(1.0, 1.0, 1.0), q=0.1, c0=10.0, c1=0.5,
                    precision="fp32")
    _check(ows, gws, 10, psi_tol=1e-5, e_tol=1e-6)
