"""Parity of the kernel instantiations that the benchmarks time, against the FP64 oracle.

The spin-6 dipolar benchmark runs the fused x/site kernel on 128-point x lines (its radix
plan, 8*4*4 with 16 values per thread, and its tile differ from the 16- and 32-point lines
of the other fused-kernel tests), and its energy evaluation runs the 13-component x pass
on 128-point lines.  Each case below makes exactly those instantiations run: 2-D (128, 16)
and 3-D (128, 8, 8) grids, FP64 and FP32-bulk, plain and zero-padded full-Q DDI, plus the
neighbouring line lengths of the same kernels (D = 13 at x sides 64 and 256, D = 5 at 128
and 256) and the 13-component x and site passes on radix-7 sides (112, 224).

Bars: the fused schedule equals the split schedule (separate x and lane-per-site passes)
bit for bit for the spin-6 tensor and DDI steps; FP64 psi agrees with the oracle to
per-component relative L2 1e-10 and every energy term to 1e-12 (relative, floor 1).  The
adaptive driver (run_simulation_yoshida!, yoshida.jl:11-98) takes the same accepted and
rejected attempts and ends at the same dt as the oracle's driver.

FP32-bulk against the FP64 oracle: the tensor step couples the FP32 rounding of the
dominant m = +6 component (relative 6e-8) into the noise components, whose norm is 1e-3 of
it, so their own relative error is about 6e-8 / 1e-3 ~ 1e-4 after a few sub-steps; the
whole spinor's relative L2 stays at a few 1e-6.  Both are bounded below.
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
FP32_TOTAL_TOL = 1e-5     # relative L2 of the whole spinor, FP32-bulk vs the FP64 oracle
FP32_COMP_TOL = 5e-4      # per component (the 1e-3 noise components carry the largest)


def _fp32_parity(a, ows):
    b = ows.psi
    tot = math.sqrt(float(np.sum(np.abs(a - b) ** 2)) / float(np.sum(np.abs(b) ** 2)))
    assert tot <= FP32_TOTAL_TOL, tot
    assert rel_l2(a, b) <= FP32_COMP_TOL
    eo = O.total_energy(ows)
    return eo


def _cextra(seed=42):
    r = np.random.default_rng(seed)
    ce = [0.0] * 11
    for k in range(2, 13, 2):
        ce[k - 2] = float(r.uniform(-1, 1))
    return ce


def _stretched_psi(og, D, seed):
    """m = +F Gaussian of width L/8 plus 1e-3 seeded noise, unit norm (the spin-6
    benchmark's initial state)."""
    psi = np.zeros((D,) + og.shape, complex)
    psi[0] = O.gaussian(og, tuple(0.125 * l for l in og.L))
    psi += 1e-3 * random_psi(og.shape, D, seed)
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    return psi


def _launches(ws):
    return {k: v[1] for k, v in ws.profile_read().items()}


def _energy_parity(ows, gws):
    eo, eg = O.energy_terms(ows), A.energy_terms(gws)
    for k in eo:
        assert abs(eg[k] - eo[k]) <= E_TOL * max(1.0, abs(eo[k])), (k, eg[k], eo[k])
    tot = O.total_energy(ows)
    assert abs(A.total_energy(gws) - tot) <= E_TOL * abs(tot)


SPIN6_SHAPES = [((128, 16), (24.0, 12.0)), ((128, 8, 8), (24.0, 12.0, 12.0))]


@pytest.mark.parametrize("padding", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", SPIN6_SHAPES, ids=["128x16", "128x8x8"])
def test_spin6_x128_yoshida_fused_split_oracle(shape, precision, padding):
    """xsite_kernel<T, 128, 13>: two Yoshida cores (12 fused launches each with DDI), fused
    == split bitwise; FP64 against the oracle (psi and every energy term)."""
    n, L = shape
    og, _ = grids(n, L)
    psi = _stretched_psi(og, 13, 42)
    omega = (1.0,) * len(n)
    kw = dict(omega=omega, c0=10.0, c_extra=_cextra(), ddi=True, cdd=1.0,
              precision=precision, padding=padding)
    ows, gf = pair(n, L, 6, psi, 1e-3, False, **kw)
    _, gs = pair(n, L, 6, psi, 1e-3, False, **kw)
    assert gf.set_xsite(True), "128-point spin-6 lines run the fused kernel"
    assert not gs.set_xsite(False)
    for ws in (gf, gs):
        ws.profile(True)
        A.yoshida_core(ws, 1e-3)
        A.yoshida_core(ws, 1e-3)
    nf, ns = _launches(gf), _launches(gs)
    assert nf["xsite"] > 0 and nf["sites"] == 0, nf
    assert ns["xsite"] == 0 and ns["sites"] > 0, ns
    a, b = gf.get_psi(), gs.get_psi()
    assert np.array_equal(a, b), float(np.max(np.abs(a - b)))
    O.yoshida_core(ows, 1e-3)
    O.yoshida_core(ows, 1e-3)
    if precision == "fp64":
        assert rel_l2(a, ows.psi) <= PSI_TOL
        _energy_parity(ows, gf)
    else:
        eo = _fp32_parity(a, ows)
        assert abs(A.total_energy(gf) - eo) <= 1e-6 * abs(eo)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_spin6_x128_itp_fused_split_oracle(precision):
    """ITP with the norm partials of the fused 128-point launch (renormalisation every
    step), tensor + full-Q DDI, 3 split steps."""
    n, L = (128, 8, 8), (24.0, 12.0, 12.0)
    og, _ = grids(n, L)
    psi = _stretched_psi(og, 13, 7)
    kw = dict(omega=(1, 1, 1), p=0.1, c0=10.0, c_extra=_cextra(3), ddi=True, cdd=1.0,
              precision=precision)
    ows, gf = pair(n, L, 6, psi, 0.002, True, **kw)
    _, gs = pair(n, L, 6, psi, 0.002, True, **kw)
    assert gf.set_xsite(True)
    gs.set_xsite(False)
    A.split_step(gf, 3)
    A.split_step(gs, 3)
    a, b = gf.get_psi(), gs.get_psi()
    assert np.array_equal(a, b), float(np.max(np.abs(a - b)))
    for _ in range(3):
        O.split_step(ows)
    if precision == "fp64":
        assert rel_l2(a, ows.psi) <= PSI_TOL
        _energy_parity(ows, gf)
    else:
        _fp32_parity(a, ows)


@pytest.mark.parametrize("nx", [64, 256])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_spin6_x64_x256_fused_split_oracle(nx, precision):
    """xsite_kernel<T, 64 | 256, 13> with tensor + DDI: one Yoshida core."""
    n, L = (nx, 8), (24.0, 12.0)
    og, _ = grids(n, L)
    psi = _stretched_psi(og, 13, 11)
    kw = dict(omega=(1, 1), c0=10.0, c_extra=_cextra(5), ddi=True, cdd=1.0, precision=precision)
    ows, gf = pair(n, L, 6, psi, 1e-3, False, **kw)
    _, gs = pair(n, L, 6, psi, 1e-3, False, **kw)
    assert gf.set_xsite(True)
    gs.set_xsite(False)
    A.yoshida_core(gf, 1e-3)
    A.yoshida_core(gs, 1e-3)
    a, b = gf.get_psi(), gs.get_psi()
    assert np.array_equal(a, b), float(np.max(np.abs(a - b)))
    O.yoshida_core(ows, 1e-3)
    if precision == "fp64":
        assert rel_l2(a, ows.psi) <= PSI_TOL
        _energy_parity(ows, gf)
    else:
        _fp32_parity(a, ows)


@pytest.mark.parametrize("nx", [128, 256])
@pytest.mark.parametrize("itp", [False, True])
def test_spin2_x128_x256_fused_split_oracle(nx, itp):
    """xsite_kernel<double, 128 | 256, 5>: F = 2 Euler spin mixing with DDI.  The spin-2
    rotation angle is formed in registers and rounds differently in the two kernels
    (relative 1e-13 allowed between them); both agree with the oracle to 1e-10."""
    n, L = (nx, 16), (20.0, 12.0)
    og, _ = grids(n, L)
    psi = random_psi(og.shape, 5, 13) * O.gaussian(og, (3.0, 2.0))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    kw = dict(omega=(1, 1), p=0.2, q=0.05, c0=8.0, c1=-0.7, ddi=True, cdd=0.8)
    ows, gf = pair(n, L, 2, psi, 0.002, itp, **kw)
    _, gs = pair(n, L, 2, psi, 0.002, itp, **kw)
    assert gf.set_xsite(True)
    gs.set_xsite(False)
    gf.profile(True)
    A.split_step(gf, 3)
    A.split_step(gs, 3)
    assert _launches(gf)["xsite"] > 0
    a, b = gf.get_psi(), gs.get_psi()
    assert rel_l2(a, b) <= 1e-13
    for _ in range(3):
        O.split_step(ows)
    assert rel_l2(a, ows.psi) <= PSI_TOL
    _energy_parity(ows, gf)


def test_spin6_x128_adaptive_yoshida_driver():
    """run_simulation_yoshida! on the 128-point spin-6 dipolar grid (tensor + full-Q DDI),
    with a save point inside the run so that one step is clamped: same accepted and
    rejected counts, same final dt, psi and snapshot energies as the oracle's driver."""
    n, L = (128, 8, 8), (24.0, 12.0, 12.0)
    og, _ = grids(n, L)
    psi = _stretched_psi(og, 13, 42)
    ows, gws = pair(n, L, 6, psi, 1e-3, False, omega=(1, 1, 1), c0=10.0, c_extra=_cextra(),
                    ddi=True, cdd=1.0)
    assert gws.set_xsite(True)
    kw = dict(t_end=0.012, save_interval=0.005)
    ro = O.run_simulation_yoshida(ows, dt_init=1e-3, tol=1e-6, **kw)
    rg = A.run_simulation_yoshida(gws, adaptive=G.AdaptiveDtParams(dt_init=1e-3, tol=1e-6), **kw)
    assert (rg["n_accepted"], rg["n_rejected"]) == (ro["n_accepted"], ro["n_rejected"])
    assert rg["n_accepted"] >= 3
    assert rg["final_dt"] == pytest.approx(ro["final_dt"], rel=1e-9)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    res = rg["result"]
    assert np.allclose(res.times, ro["result"]["times"], rtol=0, atol=1e-15)
    for eg, eo in zip(res.energies, ro["result"]["energies"]):
        assert abs(eg - eo) <= E_TOL * abs(eo)


@pytest.mark.parametrize("padding", [False, True])
def test_spin6_radix7_3d_split_step(padding):
    """13-component x pass and site pass on radix-7 sides (112 x 56 x 8), tensor + full-Q
    DDI, plain and zero-padded (224- and 112-point DDI lines)."""
    n, L = (112, 56, 8), (24.0, 16.0, 10.0)
    og, _ = grids(n, L)
    psi = _stretched_psi(og, 13, 17)
    ows, gws = pair(n, L, 6, psi, 1e-3, False, omega=(1, 1, 1), c0=10.0, c_extra=_cextra(9),
                    ddi=True, cdd=1.0, padding=padding)
    assert not gws.set_xsite(True), "non-power-of-two sides run the split passes"
    A.split_step(gws, 1)
    O.split_step(ows)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    _energy_parity(ows, gws)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_spin6_radix7_224_yoshida(precision):
    """224-point lines (radix 7, 28 values per thread) with 13 components: Yoshida core."""
    n, L = (224, 16), (28.0, 12.0)
    og, _ = grids(n, L)
    psi = _stretched_psi(og, 13, 19)
    ows, gws = pair(n, L, 6, psi, 1e-3, False, omega=(1, 1), c0=10.0, c_extra=_cextra(21),
                    ddi=True, cdd=1.0, precision=precision)
    A.yoshida_core(gws, 1e-3)
    O.yoshida_core(ows, 1e-3)
    if precision == "fp64":
        assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
        _energy_parity(ows, gws)
    else:
        _fp32_parity(gws.get_psi(), ows)
