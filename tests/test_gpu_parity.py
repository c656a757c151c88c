"""GPU (libbecgpu.so, FP64) vs the FP64 oracle on identical seeded inputs.

Bar: per-component relative L2 <= 1e-10 for psi and
relative energy <= 1e-12 for the plain FP64 path.
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


def _gs_inputs(n, L, F, state="polar"):
    og, gg = grids(n, L)
    return O.init_psi(og, F, state)


def test_cfg1_ground_state_1d_spin1():
    """1D spin-1 ITP: 64 points, box 20, c0=10, c1=-0.5, omega=1, dt=0.005, 5000 steps,
    tol 1e-10, polar start (the CPU-runnable configuration)."""
    og, gg = grids((64,), (20.0,))
    psi0 = O.init_psi(og, 1, "polar")
    ows = O.make_workspace(og, 1, psi0, 0.005, True, omega=(1.0,), c0=10.0, c1=-0.5, n_steps=5000)
    ro = O.find_ground_state(ows, 5000, 1e-10)
    rg = G.find_ground_state(grid=gg, atom=G.Rb87, interactions=G.InteractionParams(10.0, -0.5),
                             potential=G.HarmonicTrap(1.0), dt=0.005, n_steps=5000, tol=1e-10,
                             psi_init=psi0)
    print("oracle", ro, "gpu", {k: v for k, v in rg.items() if k != "workspace"})
    assert rg["converged"] == ro["converged"]
    assert rg["steps"] == ro["steps"]
    assert abs(rg["energy"] - ro["energy"]) <= E_TOL * abs(ro["energy"])
    assert rel_l2(rg["workspace"].get_psi(), ows.psi) <= PSI_TOL


@pytest.mark.parametrize("itp", [True, False])
def test_split_step_2d_spin1(itp):
    n, L = (64, 32), (16.0, 12.0)
    og, _ = grids(n, L)
    psi = O.init_psi(og, 1, "uniform") * (1 + 0.1 * random_psi(og.shape, 1, 3)[0])
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 1, psi, 0.004, itp, omega=(1.0, 1.3), p=0.3, q=0.2, c0=12.0, c1=-0.7,
                    c_lhy=0.5)
    for _ in range(40):
        O.split_step(ows)
    A.split_step(gws, 40)
    err = rel_l2(gws.get_psi(), ows.psi)
    print("err", err)
    assert err <= PSI_TOL


@pytest.mark.parametrize("F,itp", [(1, False), (1, True), (2, False), (2, True)])
def test_nematic_split_step_2d(F, itp):
    """Nematic singlet-pair step (c2 = c_extra[1], nematic.jl:34-119) with spin mixing and
    Zeeman, 2-D: psi after 30 steps and every energy term (the nematic energy is reported
    in the tensor slot, energy.jl:43-46)."""
    n, L = (32, 32), (12.0, 12.0)
    og, _ = grids(n, L)
    D = 2 * F + 1
    psi = O.gaussian(og, (2.5, 2.5)) * (0.6 + 0.3 * random_psi(og.shape, D, 17 + F))
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, F, psi, 0.003, itp, omega=(1.0, 1.1), p=0.2, q=0.1, c0=8.0, c1=-0.4,
                    c_extra=(1.7,))
    for _ in range(30):
        O.split_step(ows)
    A.split_step(gws, 30)
    err = rel_l2(gws.get_psi(), ows.psi)
    eo, eg = O.energy_terms(ows), A.energy_terms(gws)
    print("F", F, "itp", itp, "err", err, "tensor slot", eo["tensor"], eg["tensor"])
    assert err <= PSI_TOL
    assert eo["tensor"] != 0.0
    for k in ("kin", "trap", "zeeman", "c0", "c1", "tensor"):
        assert abs(eg[k] - eo[k]) <= E_TOL * max(abs(eo[k]), 1e-300), k


def test_nematic_f6_site_kernel():
    """F=6 with only c2 (no rank >= 4 channel: the tensor step stays off), so the nematic
    step runs in the lane-per-site pass."""
    n, L = (16, 16, 16), (10.0, 10.0, 10.0)
    og, _ = grids(n, L)
    psi = O.gaussian(og, (2.0, 2.0, 2.0)) * (0.5 + 0.3 * random_psi(og.shape, 13, 41))
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 6, psi, 0.002, False, omega=(1.0, 1.0, 1.0), c0=5.0, c1=0.3,
                    c_extra=(2.2,))
    assert ows.T is None
    for _ in range(6):
        O.split_step(ows)
    A.split_step(gws, 6)
    err = rel_l2(gws.get_psi(), ows.psi)
    print("err", err)
    assert err <= PSI_TOL


@pytest.mark.parametrize("F", [1, 2])
def test_loss_split_step_and_leapfrog_2d(F):
    """Dipolar relaxation + three-body loss (losses.jl:25-82): applied after V K V in real-time
    split_step! and after the merged potential step of the leapfrog (simulation.jl:74-76);
    skipped in imaginary time (split_step.jl:21-23)."""
    n, L = (32, 32), (12.0, 12.0)
    og, _ = grids(n, L)
    D = 2 * F + 1
    psi = O.gaussian(og, (2.0, 2.5)) * (0.6 + 0.3 * random_psi(og.shape, D, 5 + F))
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    loss = (0.8, 0.3)
    ows, gws = pair(n, L, F, psi, 0.004, False, omega=(1.0, 1.2), q=0.1, c0=6.0, c1=0.5,
                    loss=loss)
    for _ in range(25):
        O.split_step(ows)
    A.split_step(gws, 25)
    err = rel_l2(gws.get_psi(), ows.psi)
    print("split_step err", err, "norm", O.total_norm(ows.psi, og.dV))
    assert err <= PSI_TOL
    assert O.total_norm(ows.psi, og.dV) < 0.999
    ows, gws = pair(n, L, F, psi, 0.004, False, omega=(1.0, 1.2), q=0.1, c0=6.0, c1=0.5,
                    loss=loss, n_steps=30, save_every=10)
    ro = O.run_simulation(ows)
    rg = A.run_simulation(gws, keep_snapshots=False)
    print("leapfrog norms", ro["norms"], rg.norms)
    assert np.allclose(rg.norms, ro["norms"], rtol=1e-12, atol=0)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    ows, gws = pair(n, L, F, psi, 0.004, True, omega=(1.0, 1.2), c0=6.0, c1=0.5, loss=(100.0, 100.0))
    for _ in range(10):
        O.split_step(ows)
    A.split_step(gws, 10)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    assert abs(O.total_norm(ows.psi, og.dV) - 1.0) < 1e-12


def test_loss_yoshida_driver():
    """run_simulation_yoshida! applies the loss after each accepted step (yoshida.jl:64-66)."""
    n, L = (32, 32), (12.0, 12.0)
    og, gg = grids(n, L)
    psi = O.gaussian(og, (2.0, 2.0)) * (0.6 + 0.3 * random_psi(og.shape, 3, 8))
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 1, psi, 1e-3, False, omega=(1.0, 1.0), c0=6.0, c1=-0.5,
                    loss=(0.6, 0.1))
    ro = O.run_simulation_yoshida(ows, dt_init=1e-3, t_end=0.05, save_interval=0.02)
    rg = A.run_simulation_yoshida(gws, adaptive=G.AdaptiveDtParams(dt_init=1e-3), t_end=0.05,
                                  save_interval=0.02, keep_snapshots=False)
    print("acc", ro["n_accepted"], rg["n_accepted"], "norms", ro["result"]["norms"], rg["result"].norms)
    assert rg["n_accepted"] == ro["n_accepted"] and rg["n_rejected"] == ro["n_rejected"]
    assert np.allclose(rg["result"].norms, ro["result"]["norms"], rtol=1e-12, atol=0)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL


@pytest.mark.parametrize("F,target", [(1, 0.3), (2, -0.5)])
def test_constrained_mz_ground_state_1d(F, target):
    """find_ground_state(target_magnetization=...) (ground_state.jl:118-146,279-320)."""
    og, gg = grids((64,), (20.0,))
    psi0 = O.init_psi(og, F, "uniform")
    ows = O.make_workspace(og, F, psi0, 0.005, True, omega=(1.0,), c0=10.0, c1=-0.5, q=0.05,
                           n_steps=400, normalize_every=0)
    ro = O.find_ground_state(ows, 400, 1e-10, target_mz=target)
    rg = G.find_ground_state(grid=gg, atom=G.AtomSpecies("test", 1.0, F),
                             interactions=G.InteractionParams(10.0, -0.5),
                             zeeman=G.ZeemanParams(0.0, 0.05), potential=G.HarmonicTrap(1.0),
                             dt=0.005, n_steps=400, tol=1e-10, psi_init=psi0,
                             target_magnetization=target)
    print("oracle", ro, "gpu", {k: v for k, v in rg.items() if k != "workspace"})
    assert rg["steps"] == ro["steps"]
    assert abs(rg["energy"] - ro["energy"]) <= E_TOL * abs(ro["energy"])
    assert rel_l2(rg["workspace"].get_psi(), ows.psi) <= PSI_TOL
    assert abs(A.magnetization(rg["workspace"]) - target) < 1e-10


def test_adaptive_dt_ground_state_1d():
    """find_ground_state(adaptive_dt=true, dt_max) (ground_state.jl:172-228)."""
    og, gg = grids((64,), (15.0,))
    psi0 = O.init_psi(og, 1, "polar")
    ows = O.make_workspace(og, 1, psi0, 0.001, True, omega=(1.0,), c0=10.0, c1=-0.5, n_steps=2000)
    ro = O.find_ground_state_adaptive(ows, 2000, 1e-8, 0.01)
    rg = G.find_ground_state(grid=gg, atom=G.Rb87, interactions=G.InteractionParams(10.0, -0.5),
                             potential=G.HarmonicTrap(1.0), dt=0.001, n_steps=2000, tol=1e-8,
                             psi_init=psi0, adaptive_dt=True, dt_max=0.01)
    print("oracle", ro, "gpu", {k: v for k, v in rg.items() if k != "workspace"})
    assert rg["steps"] == ro["steps"] and rg["converged"] == ro["converged"]
    assert abs(rg["final_dt"] - ro["final_dt"]) <= 1e-15 * ro["final_dt"]
    assert abs(rg["energy"] - ro["energy"]) <= E_TOL * abs(ro["energy"])
    assert rel_l2(rg["workspace"].get_psi(), ows.psi) <= PSI_TOL


def test_leapfrog_3d_spin1_snapshots():
    n, L = (32, 32, 32), (20.0, 20.0, 20.0)
    og, _ = grids(n, L)
    psi = O.thomas_fermi_psi(og, 1, O.harmonic_potential(og, (1, 1, 1)), 10.0)
    noise = 1e-3 * random_psi(og.shape, 3, 7)
    psi[0] += noise[0]
    psi[2] += noise[2]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 1, psi, 0.001, False, omega=(1, 1, 1), q=0.1, c0=10.0, c1=0.5,
                    n_steps=20, save_every=5)
    ro = O.run_simulation(ows)
    rg = A.run_simulation(gws, keep_snapshots=False)
    assert len(rg.energies) == len(ro["energies"]) == 5
    for eg, eo in zip(rg.energies, ro["energies"]):
        assert abs(eg - eo) <= E_TOL * abs(eo)
    for ng, no in zip(rg.norms, ro["norms"]):
        assert abs(ng - no) <= 1e-12
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL


@pytest.mark.parametrize("F", [2, 6])
def test_spin_mixing_euler_3d(F):
    n, L = (16, 16, 16), (10.0, 10.0, 10.0)
    og, _ = grids(n, L)
    D = 2 * F + 1
    psi = random_psi(og.shape, D, 11 + F) * O.gaussian(og, (2.0, 2.0, 2.0))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, F, psi, 0.002, False, omega=(1, 1, 1), p=0.2, q=0.05, c0=8.0, c1=3.0)
    for _ in range(5):
        O.split_step(ows)
    A.split_step(gws, 5)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL


@pytest.mark.parametrize("itp", [False, True])
@pytest.mark.parametrize("secular", [False, True])
def test_ddi_3d_split_step(itp, secular):
    n, L = (16, 32, 16), (12.0, 14.0, 12.0)
    og, _ = grids(n, L)
    F = 1
    psi = random_psi(og.shape, 3, 21) * O.gaussian(og, (2.5, 2.5, 2.5))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, F, psi, 0.002, itp, omega=(1, 1, 1), c0=6.0, c1=-0.4, ddi=True, cdd=2.0,
                    secular=secular)
    for _ in range(4):
        O.split_step(ows)
    A.split_step(gws, 4)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    eo = O.energy_terms(ows)
    eg = A.energy_terms(gws)
    for k in ("kin", "trap", "c0", "c1", "ddi"):
        assert abs(eg[k] - eo[k]) <= 1e-12 * max(1.0, abs(eo[k])), k


def test_ddi_2d_spin6_energy():
    n, L = (32, 16), (10.0, 8.0)
    og, _ = grids(n, L)
    F = 6
    psi = random_psi(og.shape, 13, 5)
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, F, psi, 0.001, False, omega=(1, 1), p=0.5, q=0.1, c0=5.0, c1=0.3,
                    ddi=True, cdd=1.5)
    eo = O.energy_terms(ows)
    eg = A.energy_terms(gws)
    for k in eo:
        assert abs(eg[k] - eo[k]) <= 1e-12 * max(1.0, abs(eo[k])), (k, eg[k], eo[k])
    O.split_step(ows)
    A.split_step(gws, 1)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL


@pytest.mark.parametrize("itp", [False, True])
def test_ddi_padded_3d_split_step(itp):
    """Zero-padded DDI convolution (ddi_padded.jl): spectra on the (2n)^3 grid, Phi cropped;
    the energy's DDI term stays on the unpadded convolution (energy.jl:111-121)."""
    n, L = (16, 32, 16), (12.0, 14.0, 12.0)
    og, _ = grids(n, L)
    psi = random_psi(og.shape, 3, 23) * O.gaussian(og, (2.5, 2.5, 2.5))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 1, psi, 0.002, itp, omega=(1, 1, 1), c0=6.0, c1=-0.4, ddi=True, cdd=2.0,
                    padding=True)
    assert ows.Qpad is not None
    for _ in range(4):
        O.split_step(ows)
    A.split_step(gws, 4)
    err = rel_l2(gws.get_psi(), ows.psi)
    print("padded err", err)
    assert err <= PSI_TOL
    eo, eg = O.energy_terms(ows), A.energy_terms(gws)
    for k in ("kin", "trap", "c0", "c1", "ddi"):
        assert abs(eg[k] - eo[k]) <= 1e-12 * max(1.0, abs(eo[k])), k


def test_ddi_padded_spin6_2d_and_fp32():
    """Padded DDI for spin 6 on a 2-D grid (FP64 parity) and the same steps FP32-bulk."""
    n, L = (32, 16), (10.0, 8.0)
    og, _ = grids(n, L)
    psi = random_psi(og.shape, 13, 6) * O.gaussian(og, (2.0, 2.0))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 6, psi, 0.001, False, omega=(1, 1), p=0.5, q=0.1, c0=5.0, c1=0.3,
                    ddi=True, cdd=1.5, padding=True)
    _, gws32 = pair(n, L, 6, psi, 0.001, False, omega=(1, 1), p=0.5, q=0.1, c0=5.0, c1=0.3,
                    ddi=True, cdd=1.5, padding=True, precision="fp32")
    for _ in range(3):
        O.split_step(ows)
    A.split_step(gws, 3)
    A.split_step(gws32, 3)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    e32 = rel_l2(gws32.get_psi(), ows.psi)
    print("fp32 padded err", e32)
    assert e32 <= 1e-5


def test_strang_and_yoshida_core():
    n, L = (16, 16, 32), (10.0, 10.0, 12.0)
    og, _ = grids(n, L)
    psi = random_psi(og.shape, 5, 9) * O.gaussian(og, (2.0, 2.0, 2.5))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 2, psi, 0.001, False, omega=(1, 1, 1), c0=7.0, c1=0.8, ddi=True, cdd=1.0)
    O.strang_core(ows, 0.003)
    A.strang_core(gws, 0.003)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    O.yoshida_core(ows, 0.004)
    A.yoshida_core(gws, 0.004)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL


def test_adaptive_yoshida_accept_reject():
    n, L = (16, 16, 16), (10.0, 10.0, 10.0)
    og, _ = grids(n, L)
    psi = random_psi(og.shape, 3, 4) * O.gaussian(og, (2.0, 2.0, 2.0))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 1, psi, 0.001, False, omega=(1, 1, 1), c0=10.0, c1=0.5, ddi=True, cdd=1.0)
    ro = O.run_simulation_yoshida(ows, dt_init=0.01, tol=1e-6, t_end=0.05, save_interval=0.02)
    rg = A.run_simulation_yoshida(gws, adaptive=G.AdaptiveDtParams(dt_init=0.01, tol=1e-6),
                                  t_end=0.05, save_interval=0.02, keep_snapshots=False)
    assert (rg["n_accepted"], rg["n_rejected"]) == (ro["n_accepted"], ro["n_rejected"])
    assert abs(rg["final_dt"] - ro["final_dt"]) <= 1e-9 * ro["final_dt"]
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL


def test_batched_instances_match_single():
    """Three instances with different c1/p/omega in one workspace == three separate runs."""
    n, L = (32, 32), (12.0, 12.0)
    og, gg = grids(n, L)
    sp = G.SimParams(dt=0.003, n_steps=30, imaginary_time=True)
    ws = G.Workspace(grid=gg, F=1, n_instances=3, sim_params=sp)
    cfgs = [(-0.5, 0.1, 1.0), (0.3, 0.0, 1.2), (0.9, 0.4, 0.8)]
    psis = []
    for i, (c1, p, w) in enumerate(cfgs):
        ws.set_instance(i, interactions=G.InteractionParams(9.0, c1), zeeman=G.ZeemanParams(p, 0.05),
                        potential=G.HarmonicTrap(w, w))
        psi = O.init_psi(og, 1, ["polar", "ferromagnetic", "uniform"][i])
        ws.set_psi(psi, i)
        psis.append(psi)
    A.split_step(ws, 30)
    for i, (c1, p, w) in enumerate(cfgs):
        ows = O.make_workspace(og, 1, psis[i], 0.003, True, omega=(w, w), p=p, q=0.05, c0=9.0, c1=c1)
        for _ in range(30):
            O.split_step(ows)
        assert rel_l2(ws.get_psi(i), ows.psi) <= PSI_TOL


def _cfg3_extra(seed=42):
    """c_extra with even-rank couplings c2..c12 ~ U[-1, 1] (odd entries 0)."""
    r = np.random.default_rng(seed)
    ce = [0.0] * 11
    for k in range(2, 13, 2):
        ce[k - 2] = float(r.uniform(-1, 1))
    return ce


@pytest.mark.parametrize("itp", [False, True])
def test_tensor_f6_3d_split_step(itp):
    n, L = (16, 16, 16), (12.0, 12.0, 12.0)
    og, _ = grids(n, L)
    psi = random_psi(og.shape, 13, 31) * O.gaussian(og, (2.5, 2.5, 2.5))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 6, psi, 0.002, itp, omega=(1, 1, 1), p=0.1, c0=10.0,
                    c_extra=_cfg3_extra())
    assert gws.tensor and ows.T is not None
    eo = O.energy_terms(ows)
    eg = A.energy_terms(gws)
    for k in eo:
        assert abs(eg[k] - eo[k]) <= 1e-12 * max(1.0, abs(eo[k])), (k, eg[k], eo[k])
    for _ in range(2):
        O.split_step(ows)
    A.split_step(gws, 2)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL


def test_tensor_f6_ddi_yoshida_core():
    """One Yoshida attempt of the F=6 dipolar configuration (tensor + DDI) on 16^3."""
    n, L = (16, 16, 16), (24.0, 24.0, 24.0)
    og, _ = grids(n, L)
    psi = np.zeros((13,) + og.shape, complex)
    psi[0] = O.gaussian(og, (3.0, 3.0, 3.0))
    psi += 1e-3 * random_psi(og.shape, 13, 42)
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 6, psi, 1e-3, False, omega=(1, 1, 1), c0=10.0, c_extra=_cfg3_extra(),
                    ddi=True, cdd=1.0)
    O.yoshida_core(ows, 1e-3)
    A.yoshida_core(gws, 1e-3)
    assert rel_l2(gws.get_psi(), ows.psi) <= PSI_TOL
    eo, eg = O.total_energy(ows), A.total_energy(gws)
    assert abs(eg - eo) <= E_TOL * abs(eo)


# --- FP32-bulk path: FP32 psi/FFT/pointwise, FP64 reductions and phase arguments ------
def test_fp32_bulk_ground_state_energy_cfg1():
    """cfg1 ground state in FP32 bulk: relative energy error <= 1e-7 vs the FP64 oracle."""
    og, gg = grids((64,), (20.0,))
    psi0 = O.init_psi(og, 1, "polar")
    ows = O.make_workspace(og, 1, psi0, 0.005, True, omega=(1.0,), c0=10.0, c1=-0.5, n_steps=5000)
    ro = O.find_ground_state(ows, 5000, 1e-10)
    rg = G.find_ground_state(grid=gg, atom=G.Rb87, interactions=G.InteractionParams(10.0, -0.5),
                             potential=G.HarmonicTrap(1.0), dt=0.005, n_steps=5000, tol=1e-10,
                             psi_init=psi0, precision="fp32")
    rel = abs(rg["energy"] - ro["energy"]) / abs(ro["energy"])
    print("fp32 cfg1 rel energy error", rel)
    assert rel <= 1e-7
    assert rel_l2(rg["workspace"].get_psi(), ows.psi) <= 1e-4


def test_fp32_bulk_tensor_ddi_step():
    n, L = (16, 16, 16), (24.0, 24.0, 24.0)
    og, _ = grids(n, L)
    psi = np.zeros((13,) + og.shape, complex)
    psi[0] = O.gaussian(og, (3.0, 3.0, 3.0))
    psi += 1e-3 * random_psi(og.shape, 13, 42)
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 6, psi, 1e-3, False, omega=(1, 1, 1), c0=10.0, c_extra=_cfg3_extra(),
                    ddi=True, cdd=1.0, precision="fp32")
    for _ in range(3):
        O.yoshida_core(ows, 1e-3)
        A.yoshida_core(gws, 1e-3)
    err = rel_l2(gws.get_psi(), ows.psi)
    eo, eg = O.total_energy(ows), A.total_energy(gws)
    print("fp32 tensor/ddi psi err", err, "energy rel err", abs(eg - eo) / abs(eo))
    assert err <= 5e-5
    assert abs(eg - eo) <= 1e-6 * abs(eo)


def test_fp32_bulk_leapfrog_3d():
    n, L = (32, 32, 32), (20.0, 20.0, 20.0)
    og, _ = grids(n, L)
    psi = O.thomas_fermi_psi(og, 1, O.harmonic_potential(og, (1, 1, 1)), 10.0)
    psi[0] += 1e-3 * random_psi(og.shape, 1, 7)[0]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, 1, psi, 0.001, False, omega=(1, 1, 1), q=0.1, c0=10.0, c1=0.5,
                    n_steps=50, save_every=10, precision="fp32")
    ro = O.run_simulation(ows)
    rg = A.run_simulation(gws, keep_snapshots=False)
    for eg, eo in zip(rg.energies, ro["energies"]):
        assert abs(eg - eo) <= 1e-6 * abs(eo)
    assert rel_l2(gws.get_psi(), ows.psi) <= 1e-5


@pytest.mark.parametrize("n,L", [((64,), (20.0,)), ((64, 32), (20.0, 12.0)),
                                 ((32, 16, 64), (8.0, 6.0, 20.0)), ((128, 128, 128), (20.0,) * 3)])
def test_fft_gradient_and_derivative(n, L):
    """bec_fft_gradient / bec_fft_derivative (fft_utils.jl:4-44) vs the oracle on a seeded
    random field (all modes, Nyquist included) and a band-limited one with an exact gradient;
    anisotropic grids so that a swapped axis fails."""
    og, gg = grids(n, L)
    f = np.random.default_rng(5).normal(size=og.shape)
    ref = O.fft_gradient(f, og)
    got = G.fft_gradient(f, gg)
    for d in range(og.ndim):
        err = np.max(np.abs(got[d] - ref[d])) / np.max(np.abs(ref[d]))
        print(n, d, err)
        assert err < 1e-12
    d = og.ndim
    one = G.fft_partial_derivative(f, gg, d)
    assert np.max(np.abs(one - ref[d - 1])) <= 1e-12 * np.max(np.abs(ref[d - 1]))
    # band-limited: exact derivative of cos(k.r + p)
    X = [O._axis_field(og, a, og.x[a]) for a in range(og.ndim)]
    kv = [2 * math.pi * (a + 1) / og.L[a] for a in range(og.ndim)]
    ph = sum(kv[a] * X[a] for a in range(og.ndim)) + 0.3
    got = G.fft_gradient(np.cos(ph) * np.ones(og.shape), gg)
    for a in range(og.ndim):
        assert np.max(np.abs(got[a] + kv[a] * np.sin(ph))) < 1e-12 * max(kv)


@pytest.mark.parametrize("mode", ["step_change", "richardson"])
@pytest.mark.parametrize("n,L,F", [((64,), (20.0,), 1), ((32, 32), (12.0, 12.0), 2)])
def test_run_simulation_adaptive(mode, n, L, F):
    """run_simulation_adaptive! (adaptive.jl:57-296) vs the oracle: same accept/reject
    sequence, final dt, snapshots and psi."""
    og, gg = grids(n, L)
    D = 2 * F + 1
    psi = O.init_psi(og, F, "uniform") * (1 + 0.05 * np.random.default_rng(3).normal(size=og.shape))[None]
    psi = psi + 0.02 * random_psi(og.shape, D, 4)
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    omega = (1.0,) * len(n)
    ows, gws = pair(n, L, F, psi, 0.001, False, omega=omega, c0=10.0, c1=-0.5, q=0.1)
    ro = O.run_simulation_adaptive(ows, dt_init=0.002, dt_min=1e-5, dt_max=0.01, tol=1e-4,
                                   error_mode=mode, t_end=0.15, save_interval=0.05)
    rg = G.run_simulation_adaptive(gws, adaptive=G.AdaptiveDtParams(dt_init=0.002, dt_min=1e-5,
                                                                    dt_max=0.01, tol=1e-4,
                                                                    error_mode=mode),
                                   t_end=0.15, save_interval=0.05)
    print(mode, n, ro["n_accepted"], ro["n_rejected"], rg["n_accepted"], rg["n_rejected"])
    assert (rg["n_accepted"], rg["n_rejected"]) == (ro["n_accepted"], ro["n_rejected"])
    assert rg["final_dt"] == pytest.approx(ro["final_dt"], rel=1e-9)
    assert np.allclose(rg["result"].times, ro["result"]["times"], rtol=0, atol=1e-12)
    for a, b in zip(rg["result"].energies, ro["result"]["energies"]):
        assert abs(a - b) <= E_TOL * abs(b)
    got = gws.get_psi(0)
    for c in range(D):
        assert rel_l2(got[c], ows.psi[c]) <= PSI_TOL


@pytest.mark.parametrize("nx", [8, 16, 32])
@pytest.mark.parametrize("F,itp", [(1, False), (1, True), (2, False)])
def test_short_1d_lines(nx, F, itp):
    """1-D lines of 16 and 32 points run x-pass blocks with fewer threads than op slots
    (6 or 12 for spin 1): every operator before and after the kinetic step must still run."""
    n, L = (nx,), (10.0,)
    og, gg = grids(n, L)
    D = 2 * F + 1
    psi = O.init_psi(og, F, "uniform") + 0.05 * random_psi(og.shape, D, 11)
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    ows, gws = pair(n, L, F, psi, 0.01, itp, omega=(1.0,), c0=1.0, c1=-0.3, q=0.2,
                    n_steps=40, save_every=20)
    ro = O.run_simulation(ows)
    rg = A.run_simulation(gws, keep_snapshots=False)
    got = gws.get_psi(0)
    for c in range(D):
        assert rel_l2(got[c], ows.psi[c]) <= PSI_TOL
    for a, b in zip(rg.energies, ro["energies"]):
        assert abs(a - b) <= E_TOL * abs(b)
