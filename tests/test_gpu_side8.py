"""Grids with 8-point sides against the FP64 oracle.

The reference's own tests build 8-point grids (test/test_ddi.jl:169,203,209 use (8, 8, 8) and
(8,) with DDI, test/test_nematic_tensor.jl:77 and test/test_potentials.jl:12 use (8, 8)).
An 8-point line is one radix-8 butterfly per thread (one thread per line), the DDI half
spectrum has 5 bins, and the fused x/site kernel does not take such tiles (its lines are
16..256 points), so these steps run the x and site passes.  Mixed sides (8 with 16 or 32)
make a swapped axis fail.
"""
import math

import numpy as np
import pytest

from oracle import spinor_oracle as O
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


def _check(ows, gws, nsteps):
    for _ in range(nsteps):
        O.split_step(ows)
    A.split_step(gws, nsteps)
    got = gws.get_psi()
    for c in range(got.shape[0]):
        assert rel_l2(got[c], ows.psi[c]) <= PSI_TOL, c
    eo = O.total_energy(ows)
    assert abs(A.total_energy(gws) - eo) <= E_TOL * abs(eo)


@pytest.mark.parametrize("itp", [False, True])
@pytest.mark.parametrize("n,L,F", [((8,), (6.0,), 1), ((8,), (6.0,), 2), ((8, 8), (6.0, 6.0), 1),
                                   ((8, 16), (6.0, 9.0), 2), ((8, 8, 8), (6.0,) * 3, 1),
                                   ((16, 8, 32), (8.0, 6.0, 12.0), 2)])
def test_side8_split_step(n, L, F, itp):
    psi = _psi(n, L, F, 3)
    ows, gws = pair(n, L, F, psi, 0.005, itp, omega=(1.0,) * len(n), p=0.1, q=0.2, c0=8.0,
                    c1=-0.4, c_lhy=0.3)
    _check(ows, gws, 20)


@pytest.mark.parametrize("itp", [False, True])
def test_side8_spin6_tensor(itp):
    n, L = (8, 8, 16), (6.0, 6.0, 9.0)
    psi = _psi(n, L, 6, 7)
    ows, gws = pair(n, L, 6, psi, 0.002, itp, omega=(1.0, 1.0, 1.0), q=0.05, c0=10.0,
                    c_extra=_cextra(5))
    _check(ows, gws, 10)


@pytest.mark.parametrize("padding", [False, True])
@pytest.mark.parametrize("secular", [False, True])
@pytest.mark.parametrize("n,L", [((8, 8, 8), (10.0,) * 3), ((8,), (10.0,)), ((8, 16), (8.0, 10.0))])
def test_side8_ddi(n, L, secular, padding):
    """DDI on the reference's 8^3 and (8,) grids (test_ddi.jl:169,209): full Q_ab and secular
    (k_z = 0 below three dimensions, ddi.jl:19-29), plain and zero-padded spectra (5 and 9
    half-spectrum bins along x).  On the 1-D grid the Q_ab product is its own launch."""
    psi = _psi(n, L, 1, 11)
    ows, gws = pair(n, L, 1, psi, 0.002, False, omega=(1.0,) * len(n), c0=10.0, c1=-0.5,
                    ddi=True, cdd=2.0, secular=secular, padding=padding)
    _check(ows, gws, 10)


def test_side8_yoshida_spin2_ddi():
    n, L = (8, 8, 8), (8.0,) * 3
    psi = _psi(n, L, 2, 13)
    ows, gws = pair(n, L, 2, psi, 1e-3, False, omega=(1.0, 1.0, 1.0), c0=8.0, c1=-0.7,
                    ddi=True, cdd=1.0)
    for _ in range(3):
        O.yoshida_core(ows, 1e-3)
        A.yoshida_core(gws, 1e-3)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    eo = O.total_energy(ows)
    assert abs(A.total_energy(gws) - eo) <= E_TOL * abs(eo)


def test_side8_fp32_spin6_ddi():
    """FP32 bulk on an 8^3 spin-6 tensor + DDI grid: error stated against the FP64 oracle."""
    n, L = (8, 8, 8), (8.0,) * 3
    psi = _psi(n, L, 6, 17)
    ows, gws = pair(n, L, 6, psi, 1e-3, False, omega=(1.0, 1.0, 1.0), c0=10.0,
                    c_extra=_cextra(9), ddi=True, cdd=1.0, precision="fp32")
    for _ in range(10):
        O.split_step(ows)
    A.split_step(gws, 10)
    err = rel_l2(gws.get_psi(), ows.psi)
    print("fp32 rel L2", err)
    assert err <= 1e-5
    eo = O.total_energy(ows)
    assert abs(A.total_energy(gws) - eo) <= 1e-6 * abs(eo)


@pytest.mark.parametrize("n,F,itp,padding", [((64,), 1, False, False), ((64,), 2, True, False),
                                             ((256,), 1, True, True), ((128,), 2, False, True)])
def test_ddi_1d_lines(n, F, itp, padding):
    """1-D DDI (ddi.jl:51-80 with N = 1, ddi_padded.jl on a 2n line): x-pass spectra, the
    Q_ab product launch, the potential back through the x pass; energy with its DDI term."""
    L = (24.0,)
    psi = _psi(n, L, F, 19)
    ows, gws = pair(n, L, F, psi, 0.002, itp, omega=(1.0,), q=0.1, c0=10.0, c1=-0.5,
                    ddi=True, cdd=3.0, padding=padding)
    _check(ows, gws, 20)
    assert abs(O.energy_terms(ows)["ddi"]) > 1e-6   # the term is exercised


def test_ddi_1d_fp32():
    n, L = (128,), (24.0,)
    psi = _psi(n, L, 1, 23)
    ows, gws = pair(n, L, 1, psi, 0.002, False, omega=(1.0,), c0=10.0, c1=-0.5, ddi=True,
                    cdd=3.0, precision="fp32")
    for _ in range(20):
        O.split_step(ows)
    A.split_step(gws, 20)
    err = rel_l2(gws.get_psi(), ows.psi)
    print("fp32 rel L2", err)
    assert err <= 1e-5
