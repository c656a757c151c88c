"""Raman coupling (raman.jl:8-31, energy.jl:123-135) on the GPU against the FP64 oracle.

The device applies exp(-i H_R dt) as the Euler rotation by phi = (Omega cos k.r,
-Omega sin k.r, delta); the oracle diagonalises the per-site H_R with LAPACK, as the
reference does.  F = 1 runs it in the lane-per-site kernel, F = 2 and 6 in the fused
x/site kernel.  Bar: psi relative L2 <= 1e-10, energies <= 1e-12 relative (FP64).
"""
import math

import numpy as np
import pytest

from oracle import spinor_oracle as O
from becgpu import api as A
from helpers import grids, pair, random_psi, rel_l2

pytestmark = pytest.mark.gpu


def _psi(og, D, seed, sig):
    psi = random_psi(og.shape, D, seed) * O.gaussian(og, sig)[None]
    return psi / math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)


def _check(ows, gws, steps, psi_tol=1e-10, e_tol=1e-12):
    for _ in range(steps):
        O.split_step(ows)
    A.split_step(gws, steps)
    err = rel_l2(gws.get_psi(), ows.psi)
    assert err <= psi_tol, err
    eo, eg = O.energy_terms(ows), A.energy_terms(gws)
    assert eo["raman"] != 0.0
    for k in eo:
        assert abs(eg[k] - eo[k]) <= e_tol * max(1.0, abs(eo[k])), (k, eg[k], eo[k])


@pytest.mark.parametrize("itp", [False, True])
def test_raman_spin1_2d(itp):
    n, L = (32, 32), (12.0, 10.0)
    og, _ = grids(n, L)
    ows, gws = pair(n, L, 1, _psi(og, 3, 3, (2.5, 2.0)), 0.002, itp, omega=(1, 1), p=0.1, q=0.05,
                    c0=10.0, c1=-0.5, raman=(1.5, 0.3, (2 * math.pi / 4.0, -0.7)))
    _check(ows, gws, 3)


def test_raman_spin1_1d_rabi():
    """1-D grid, k_eff = 0, delta = 0 from m = +1 (test_raman.jl 'Rabi oscillation'): the
    populations follow cos^4(b/2), sin^2(b)/2, sin^4(b/2) with b = Omega t."""
    n, L = (32,), (10.0,)
    og, _ = grids(n, L)
    psi = np.zeros((3,) + og.shape, complex)
    psi[0] = 1.0 / math.sqrt(32 * og.dV)
    ows, gws = pair(n, L, 1, psi, 0.01, False, raman=(1.0, 0.0, (0.0,)))
    A.split_step(gws, 100)
    out = gws.get_psi()
    b = 1.0 * 0.01 * 100
    pops = [float(np.sum(np.abs(out[c]) ** 2) * og.dV) for c in range(3)]
    want = [math.cos(b / 2) ** 4, math.sin(b) ** 2 / 2, math.sin(b / 2) ** 4]
    assert np.allclose(pops, want, atol=1e-12), (pops, want)


@pytest.mark.parametrize("itp", [False, True])
def test_raman_spin2_mixing_ddi_fused(itp):
    n, L = (64, 32), (20.0, 12.0)
    og, _ = grids(n, L)
    ows, gws = pair(n, L, 2, _psi(og, 5, 11, (3.0, 2.0)), 0.002, itp, omega=(1, 1), p=0.2, q=0.05,
                    c0=8.0, c1=-0.7, ddi=True, cdd=0.8, raman=(0.9, -0.2, (1.3, 0.4)))
    assert gws.set_xsite(True)
    _check(ows, gws, 3)


def test_raman_spin6_tensor_3d_fused():
    n, L = (16, 16, 16), (12.0, 12.0, 12.0)
    og, _ = grids(n, L)
    r = np.random.default_rng(42)
    ce = [0.0] * 11
    for k in range(2, 13, 2):
        ce[k - 2] = float(r.uniform(-1, 1))
    ows, gws = pair(n, L, 6, _psi(og, 13, 31, (2.5, 2.5, 2.5)), 0.002, False, omega=(1, 1, 1),
                    p=0.1, c0=10.0, c_extra=ce, raman=(0.6, 0.25, (0.8, 0.0, -1.1)))
    _check(ows, gws, 2)


def test_raman_fp32_bulk_spin2():
    n, L = (64, 32), (20.0, 12.0)
    og, _ = grids(n, L)
    ows, gws = pair(n, L, 2, _psi(og, 5, 5, (3.0, 2.0)), 0.002, False, omega=(1, 1), c0=8.0,
                    c1=-0.7, raman=(0.9, -0.2, (1.3, 0.4)), precision="fp32")
    for _ in range(5):
        O.split_step(ows)
    A.split_step(gws, 5)
    err = rel_l2(gws.get_psi(), ows.psi)
    assert err <= 1e-5, err   # FP32 bulk, as the other FP32 split-step tests
    eo, eg = O.total_energy(ows), A.total_energy(gws)
    assert abs(eg - eo) <= 1e-6 * abs(eo)


def test_raman_removed_and_params_kept():
    """set_raman(None) turns the step off; bec_set_params keeps the coupling."""
    n, L = (32, 32), (12.0, 10.0)
    og, _ = grids(n, L)
    psi = _psi(og, 3, 9, (2.5, 2.0))
    _, g1 = pair(n, L, 1, psi, 0.002, False, c0=10.0, raman=(1.5, 0.3, (1.0, 0.0)))
    _, g2 = pair(n, L, 1, psi, 0.002, False, c0=10.0)
    g1.set_raman(0, None)
    A.split_step(g1, 2)
    A.split_step(g2, 2)
    assert np.array_equal(g1.get_psi(), g2.get_psi())
    ows, g3 = pair(n, L, 1, psi, 0.002, False, c0=10.0, raman=(1.5, 0.3, (1.0, 0.0)))
    g3.set_instance(0, interactions=g3.interactions[0], zeeman=g3.zeeman[0],
                    potential=g3.potential, c_dd=0.0)
    _check(ows, g3, 2)
