"""Fused x/site kernel (x FFTs, DDI x lines and the heavy per-site operators of F = 2 and 6
in one launch) against the split schedule (separate x passes and lane-per-site passes).

Both schedules run the same device functions for every operator, FFT and DDI line, on the
same tiles.  For the spin-6 tensor and DDI steps and for diagonal steps psi agrees bit for
bit.  F = 2 spin mixing is the exception: its rotation angle c1 <F> is formed in registers
from the spinor, and the compiler rounds that chain differently in the two kernels
(measured 7e-16 absolute after one step, psi ~ 0.1), so the test allows relative 1e-13.
Parity with the FP64 oracle is checked as well, to the plain-path bar.
"""
import math

import numpy as np
import pytest

from oracle import spinor_oracle as O
from becgpu import api as A
from helpers import grids, pair, random_psi, rel_l2

pytestmark = pytest.mark.gpu


def _cextra(seed=42):
    r = np.random.default_rng(seed)
    ce = [0.0] * 11
    for k in range(2, 13, 2):
        ce[k - 2] = float(r.uniform(-1, 1))
    return ce


def _stretched_psi(og, D, seed):
    psi = np.zeros((D,) + og.shape, complex)
    psi[0] = O.gaussian(og, tuple(0.125 * l for l in og.L))
    psi += 1e-3 * random_psi(og.shape, D, seed)
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    return psi


def _both(n, L, F, psi, dt, itp, **kw):
    """Oracle workspace, fused GPU workspace, split GPU workspace (identical inputs)."""
    ows, gf = pair(n, L, F, psi, dt, itp, **kw)
    _, gs = pair(n, L, F, psi, dt, itp, **kw)
    assert gf.set_xsite(True), "the fused kernel should cover this shape"
    assert not gs.set_xsite(False)
    return ows, gf, gs


def _launches(ws):
    """{kernel class: launches} since profiling was enabled (reading resets the record)."""
    return {k: v[1] for k, v in ws.profile_read().items()}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("padding", [False, True])
def test_spin6_tensor_ddi_yoshida_fused_equals_split(precision, padding):
    n, L = (32, 32, 16), (24.0, 24.0, 12.0)
    og, _ = grids(n, L)
    psi = _stretched_psi(og, 13, 42)
    ows, gf, gs = _both(n, L, 6, psi, 1e-3, False, omega=(1, 1, 1), c0=10.0, c_extra=_cextra(),
                        ddi=True, cdd=1.0, precision=precision, padding=padding)
    for ws in (gf, gs):
        ws.profile(True)
        A.yoshida_core(ws, 1e-3)
        A.yoshida_core(ws, 1e-3)
    nf, ns = _launches(gf), _launches(gs)
    assert nf["xsite"] > 0 and nf["sites"] == 0, nf
    assert ns["xsite"] == 0 and ns["sites"] > 0, ns
    a, b = gf.get_psi(), gs.get_psi()
    assert np.array_equal(a, b), float(np.max(np.abs(a - b)))
    if not padding:   # max |Phi| partials (the padded convolution leaves them to the energy)
        assert np.array_equal(gf.itp_guard(), gs.itp_guard())
    if precision == "fp64":
        O.yoshida_core(ows, 1e-3)
        O.yoshida_core(ows, 1e-3)
        assert rel_l2(a, ows.psi) <= 1e-10
        eo = O.total_energy(ows)
        assert abs(A.total_energy(gf) - eo) <= 1e-12 * abs(eo)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_spin6_tensor_itp_fused_equals_split(precision):
    """ITP with renormalisation every step (norm partials from the fused launch)."""
    n, L = (16, 16, 64), (12.0, 12.0, 24.0)
    og, _ = grids(n, L)
    psi = random_psi(og.shape, 13, 31) * O.gaussian(og, (2.5, 2.5, 4.0))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gf, gs = _both(n, L, 6, psi, 0.002, True, omega=(1, 1, 1), p=0.1, c0=10.0,
                        c_extra=_cextra(7), precision=precision)
    A.split_step(gf, 3)
    A.split_step(gs, 3)
    a, b = gf.get_psi(), gs.get_psi()
    assert np.array_equal(a, b), float(np.max(np.abs(a - b)))
    if precision == "fp64":
        for _ in range(3):
            O.split_step(ows)
        assert rel_l2(a, ows.psi) <= 1e-10


@pytest.mark.parametrize("itp", [False, True])
def test_spin2_mixing_ddi_fused_equals_split(itp):
    """F = 2: Euler spin mixing and the DDI rotation in the fused kernel (2-D, 256-site tiles)."""
    n, L = (64, 32), (20.0, 12.0)
    og, _ = grids(n, L)
    psi = random_psi(og.shape, 5, 11) * O.gaussian(og, (3.0, 2.0))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gf, gs = _both(n, L, 2, psi, 0.002, itp, omega=(1, 1), p=0.2, q=0.05, c0=8.0, c1=-0.7,
                        ddi=True, cdd=0.8)
    A.split_step(gf, 3)
    A.split_step(gs, 3)
    a, b = gf.get_psi(), gs.get_psi()
    assert rel_l2(a, b) <= 1e-13
    # ITP overflow guard (ground_state.jl:4-29): norm and max |Phi| of the last convolution,
    # the latter from the fused kernel's own partials
    qf, qs = gf.itp_guard(), gs.itp_guard()
    assert qf[0, 1] > 0.0
    assert np.allclose(qf, qs, rtol=1e-12, atol=0)
    for _ in range(3):
        O.split_step(ows)
    assert rel_l2(a, ows.psi) <= 1e-10
    phi = ows._last_phi   # the oracle's last convolution
    assert qf[0, 1] == pytest.approx(max(float(np.max(np.abs(f))) for f in phi), rel=1e-9)


def test_fused_graph_replays_match_eager():
    """Captured split steps that run the fused kernel replay to the eager result."""
    n, L = (32, 16, 16), (24.0, 12.0, 12.0)
    og, _ = grids(n, L)
    psi = _stretched_psi(og, 13, 5)
    _, ge = pair(n, L, 6, psi, 1e-3, False, omega=(1, 1, 1), c0=10.0, c_extra=_cextra(),
                 ddi=True, cdd=1.0, precision="fp32")
    _, gg = pair(n, L, 6, psi, 1e-3, False, omega=(1, 1, 1), c0=10.0, c_extra=_cextra(),
                 ddi=True, cdd=1.0, precision="fp32")
    assert ge.set_xsite(True) and gg.set_xsite(True)
    for rep in range(3):
        A.split_step(ge, 2)
        A.split_step(gg, 2, graph=True)
        assert np.array_equal(ge.get_psi(), gg.get_psi()), rep



@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotuner_keeps_fused_tile(precision):
    """The tile autotuner times the x pass; shapes that run the fused x/site kernel keep
    the default tile instead (the x pass's timing once picked 16 rows of 32 points for the
    spin-2 shape below, a tile the fused kernel cannot take, so the shape lost the fused
    kernel for the rest of the process)."""
    import becgpu as G
    ce = (0.3, 0.0, -0.2, 0.0, 0.1, 0.0, 0.05, 0.0, -0.1, 0.0, 0.2)
    for n, F, rows in (((32, 16), 2, 8), ((64, 32, 16), 6, 1), ((128, 8, 8), 6, 1)):
        grid = G.make_grid(G.GridConfig(n, tuple(10.0 for _ in n)))
        psi = random_psi(grid.shape, 2 * F + 1, seed=3, scale=0.2)
        ws = G.make_workspace(grid=grid, atom=G.AtomSpecies("t", 1.0, F),
                              interactions=G.InteractionParams(
                                  5.0, -0.3 if F == 2 else 0.0, 0.0, ce if F == 6 else ()),
                              potential=G.HarmonicTrap(*([1.0] * len(n))),
                              sim_params=G.SimParams(dt=1e-3, n_steps=10), psi_init=psi,
                              precision=precision)          # autotunes (the default)
        assert ws.set_xsite(True), n
        assert ws.autotune() == rows, n
        ws.close()
