"""Every spin F = 0 .. 6 on the GPU against the FP64 oracle.

The reference takes any F >= 0 (types.jl:42-46): the scalar gas (F = 0), Rodrigues spin mixing
for F = 1 and Euler-angle spin mixing for larger F (spin_mixing.jl:1-66,
spinor_utils.jl:82-166), and the general-F tensor step whenever a coupling c_k with even
4 <= k <= 2F is set (initialization.jl:164-185, tensor_interaction.jl:130-240).  Cases:

* spin mixing, Zeeman, trap and full-Q DDI for F = 0, 3, 4, 5 in 1-D, 2-D and 3-D, real and
  imaginary time, FP64 against the oracle (per-component relative L2 1e-10, energy terms
  1e-12), FP32-bulk against the oracle (whole-spinor relative L2 1e-5);
* the reference's own F = 2 tensor workspace (test_tensor_interaction.jl:138-163:
  16 points, box 10, c_extra = [5, 0, 3], five split steps, norm kept to 1e-4), and the
  F = 2 tensor step with DDI in the fused x/site kernel (fused == split bit for bit);
* the tensor step for F = 3, 4, 5 (1-D, against the oracle);
* F = 1 tensor step from g_S = _c0c1_to_gS(1, 0, c1) against spin mixing
  (test_analytical_validation.jl:197-240: energies 1e-10, ten steps of propagation 1e-4),
  and against the oracle's tensor step (1e-10).
"""
import math

import numpy as np
import pytest

from oracle import spinor_oracle as O
import becgpu as G
from becgpu import api as A
from helpers import grids, pair, random_psi, rel_l2

pytestmark = pytest.mark.gpu

PSI_TOL, E_TOL = 1e-10, 1e-12


def _psi(og, D, seed, width):
    psi = random_psi(og.shape, D, seed) * O.gaussian(og, (width,) * og.ndim)[None]
    return psi / math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)


def _energy_parity(ows, gws):
    eo, eg = O.energy_terms(ows), A.energy_terms(gws)
    for k in eo:
        assert abs(eg[k] - eo[k]) <= E_TOL * max(1.0, abs(eo[k])), (k, eg[k], eo[k])


def _total_rel(a, b):
    return math.sqrt(float(np.sum(np.abs(a - b) ** 2)) / float(np.sum(np.abs(b) ** 2)))


SHAPES = [((32,), (10.0,)), ((16, 16), (8.0, 8.0)), ((16, 8, 8), (8.0, 6.0, 6.0))]


@pytest.mark.parametrize("F", [0, 3, 4, 5])
@pytest.mark.parametrize("itp", [False, True])
@pytest.mark.parametrize("shape", SHAPES, ids=["1d", "2d", "3d"])
def test_spin_mixing_ddi_general_f(F, itp, shape):
    n, L = shape
    og, _ = grids(n, L)
    D = 2 * F + 1
    psi = _psi(og, D, 100 + F, 1.5)
    kw = dict(omega=(1.0,) * len(n), p=0.3, q=0.05, c0=8.0, c1=-0.6, ddi=True, cdd=0.7)
    ows, gws = pair(n, L, F, psi, 0.002, itp, **kw)
    for _ in range(3):
        O.split_step(ows)
    A.split_step(gws, 3)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    _energy_parity(ows, gws)


@pytest.mark.parametrize("F", [0, 3, 5])
def test_spin_mixing_general_f_fp32(F):
    n, L = (16, 8, 8), (8.0, 6.0, 6.0)
    og, _ = grids(n, L)
    psi = _psi(og, 2 * F + 1, 7, 1.5)
    ows, gws = pair(n, L, F, psi, 0.002, False, omega=(1, 1, 1), p=0.3, c0=8.0, c1=-0.6,
                    ddi=True, cdd=0.7, precision="fp32")
    for _ in range(3):
        O.split_step(ows)
    A.split_step(gws, 3)
    assert _total_rel(gws.get_psi(), ows.psi) <= 1e-5


def test_reference_f2_tensor_workspace():
    """test_tensor_interaction.jl:138-163 with the oracle alongside: F = 2, c0 = c1 = 0,
    c_extra = [5, 0, 3] activates the tensor step; five real-time split steps."""
    n, L = (16,), (10.0,)
    og, gg = grids(n, L)
    psi = O.init_psi(og, 2, "polar")
    ce = (5.0, 0.0, 3.0)
    ows, gws = pair(n, L, 2, psi, 0.001, False, c_extra=ce, n_steps=10)
    assert gws.tensor
    n0 = A.total_norm(gws)
    A.split_step(gws, 5)
    for _ in range(5):
        O.split_step(ows)
    assert A.total_norm(gws) == pytest.approx(n0, rel=1e-4)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    _energy_parity(ows, gws)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("itp", [False, True])
def test_f2_tensor_ddi_fused_equals_split(precision, itp):
    """F = 2 tensor step (c_extra with k = 4) and DDI in the fused x/site kernel."""
    n, L = (32, 16), (10.0, 8.0)
    og, _ = grids(n, L)
    psi = _psi(og, 5, 21, 1.5)
    kw = dict(omega=(1, 1), p=0.2, c0=6.0, c1=-0.4, c_extra=(0.5, 0.0, 1.5), ddi=True,
              cdd=0.6, precision=precision)
    ows, gf = pair(n, L, 2, psi, 0.002, itp, **kw)
    _, gs = pair(n, L, 2, psi, 0.002, itp, **kw)
    assert gf.tensor and gf.set_xsite(True)
    gs.set_xsite(False)
    gf.profile(True)
    A.split_step(gf, 3)
    A.split_step(gs, 3)
    assert {k: v[1] for k, v in gf.profile_read().items()}["xsite"] > 0
    a, b = gf.get_psi(), gs.get_psi()
    assert np.array_equal(a, b), float(np.max(np.abs(a - b)))
    for _ in range(3):
        O.split_step(ows)
    if precision == "fp64":
        assert rel_l2(a, ows.psi) <= PSI_TOL
        _energy_parity(ows, gf)
    else:
        assert _total_rel(a, ows.psi) <= 1e-5


@pytest.mark.parametrize("F,ce", [(3, (0.4, 0.0, -0.7, 0.0, 0.9)),
                                  (4, (0.3, 0.0, 0.6, 0.0, -0.5, 0.0, 0.8)),
                                  (5, (0.2, 0.0, -0.4, 0.0, 0.5, 0.0, 0.3, 0.0, -0.6))])
@pytest.mark.parametrize("itp", [False, True])
def test_tensor_step_f3_f4_f5(F, ce, itp):
    n, L = (32,), (10.0,)
    og, _ = grids(n, L)
    psi = _psi(og, 2 * F + 1, 30 + F, 1.5)
    ows, gws = pair(n, L, F, psi, 0.002, itp, omega=(1.0,), p=0.1, c0=5.0, c1=0.3, c_extra=ce)
    assert gws.tensor and ows.T is not None
    for _ in range(3):
        O.split_step(ows)
    A.split_step(gws, 3)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    _energy_parity(ows, gws)


def _f1_tensor_ws(gg, psi, c0, c1, dt):
    """A spin-1 workspace with an explicit tensor cache g_S = _c0c1_to_gS(1, c0, c1)
    (_make_tensor_cache_from_channels); the diagonal step then carries no contact term."""
    ws = G.Workspace(grid=gg, F=1, n_instances=1, sim_params=G.SimParams(dt=dt, n_steps=10),
                     tensor=True)
    ws.set_instance(0, interactions=G.InteractionParams(0.0, 0.0))
    ws.set_tensor(0, G.physics.c0c1_to_gS(1, c0, c1))
    ws.set_psi(psi, 0)
    return ws


def test_f1_tensor_equals_spin_mixing():
    """test_analytical_validation.jl:197-240 (64 points, box 10, c0 = 100, c1 = 10, random
    spinor): tensor energy == c0 + c1 energies to 1e-10; ten real-time steps of dt = 1e-3
    with c0 = 0 agree with spin mixing to 1e-4; the tensor path equals the oracle's."""
    n, L = (64,), (10.0,)
    og, gg = grids(n, L)
    psi = random_psi(og.shape, 3, 42)
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    c0, c1 = 100.0, 10.0
    _, gsm = pair(n, L, 1, psi, 1e-3, False, c0=c0, c1=c1)
    gt = _f1_tensor_ws(gg, psi, c0, c1, 1e-3)
    es, et = A.energy_terms(gsm), A.energy_terms(gt)
    assert et["tensor"] == pytest.approx(es["c0"] + es["c1"], rel=1e-10)
    # propagation: the kinetic step is the same linear map on both sides, so whole split
    # steps with c0 = 0 compare the two potential steps
    _, gsm0 = pair(n, L, 1, psi, 1e-3, False, c1=c1)
    gt0 = _f1_tensor_ws(gg, psi, 0.0, c1, 1e-3)
    A.split_step(gsm0, 10)
    A.split_step(gt0, 10)
    a, b = gt0.get_psi(), gsm0.get_psi()
    assert np.allclose(a, b, rtol=1e-4, atol=1e-4 * float(np.max(np.abs(b))))
    ows = O.make_workspace(og, 1, psi, 1e-3, False, channels=O.c0c1_to_gS(1, 0.0, c1))
    for _ in range(10):
        O.split_step(ows)
    assert rel_l2(a, ows.psi) <= PSI_TOL


@pytest.mark.parametrize("n,L", [((8,), (4.0,)), ((16,), (6.0,)), ((16, 8), (6.0, 4.0))])
def test_scalar_gas_short_lines_and_padded_ddi(n, L):
    """F = 0 on short x lines (one line per x-pass block of n/8 threads) with zero-padded
    DDI switched on: the scalar gas has no spin density, so the dipolar step is the identity
    (the library leaves the convolution out); real and imaginary time against the oracle."""
    og, _ = grids(n, L)
    psi = _psi(og, 1, 3, 1.0)
    for itp in (False, True):
        ows, gws = pair(n, L, 0, psi, 0.002, itp, omega=(1.0,) * len(n), c0=4.0, ddi=True,
                        cdd=0.5, padding=True)
        for _ in range(3):
            O.split_step(ows)
        A.split_step(gws, 3)
        assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
        _energy_parity(ows, gws)
