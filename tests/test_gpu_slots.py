"""Device-resident psi slots (bec_copy_slot, bec_swap_slots): the adaptive drivers keep
psi_old and psi_strang on the device (yoshida.jl:36-44 copies them on the host side)."""
import numpy as np
import pytest

from becgpu import api as A
from helpers import grids, pair, random_psi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_copy_and_swap_slots(precision):
    og, _ = grids((32, 16), (10.0, 8.0))
    psi = random_psi(og.shape, 3, seed=4, scale=0.2)
    _, gws = pair((32, 16), (10.0, 8.0), 1, psi, 0.002, False, omega=(1.0, 1.0), c0=5.0,
                  c1=0.3, precision=precision)
    a = gws.get_psi()
    gws.copy_slot(1, 0)
    A.split_step(gws, 3)
    b = gws.get_psi()
    assert not np.array_equal(a, b)
    gws.swap_slots(0, 1)            # slot 0 <- the copy of a, slot 1 <- b
    assert np.array_equal(gws.get_psi(), a)
    gws.copy_slot(0, 1)
    assert np.array_equal(gws.get_psi(), b)
    gws.swap_slots(1, 1)
    assert np.array_equal(gws.get_psi(), b)
    gws.close()


def test_yoshida_driver_swap_equals_copy(monkeypatch):
    """run_simulation_yoshida moves psi_old and psi_strang with slot swaps and starts both
    cores from slot 2 when every instance is active: the same bits as the driver with
    psi copies and in-place cores."""
    og, _ = grids((32, 32), (10.0, 10.0))
    psi = random_psi(og.shape, 3, seed=9, scale=0.2)
    outs = []
    for use_copy in (False, True):
        _, gws = pair((32, 32), (10.0, 10.0), 1, psi, 0.002, False, omega=(1.0, 1.0), c0=5.0,
                      c1=-0.3)
        if use_copy:
            gws.swap_slots = lambda a, b: gws.copy_slot(a, b)
            for name in ("strang_core", "yoshida_core"):
                core = getattr(A, name)

                def from_copy(ws, dt, src_slot=0, _core=core):
                    if src_slot:
                        ws.copy_slot(0, src_slot)
                    _core(ws, dt)
                monkeypatch.setattr(A, name, from_copy)
        r = A.run_simulation_yoshida(gws, t_end=0.05, save_interval=0.025,
                                     keep_snapshots=False)
        outs.append((gws.get_psi(), r["n_accepted"], r["n_rejected"]))
        gws.close()
        monkeypatch.undo()
    assert np.array_equal(outs[0][0], outs[1][0]) and outs[0][1:] == outs[1][1:]


@pytest.mark.parametrize("case", ["spin1_2d", "spin1_1d", "spin6_3d_ddi", "spin6_3d_ddi_split",
                                  "spin2_2d_fp32"])
def test_cores_from_slot(case):
    """bec_strang_core_from / bec_yoshida_core_from equal a copy into slot 0 followed by the
    in-place core, bit for bit, whatever kernel the core starts with (x pass, fused x/site
    pass, or the split passes with the fused kernel off), and leave the source slot as it
    was."""
    ce = (0.3, 0.0, -0.2, 0.0, 0.1, 0.0, 0.05, 0.0, -0.1, 0.0, 0.2)
    kw = dict(spin1_2d=(((32, 32), (10.0, 10.0)), 1, dict(omega=(1.0, 1.0), c0=5.0, c1=-0.3)),
              spin1_1d=(((64,), (10.0,)), 1, dict(omega=(1.0,), c0=5.0, c1=-0.3, q=0.1)),
              spin6_3d_ddi=(((16, 16, 16), (8.0, 8.0, 8.0)), 6,
                            dict(omega=(1.0, 1.0, 1.0), c0=5.0, ddi=True, cdd=0.5,
                                 c_extra=ce)),
              spin6_3d_ddi_split=(((16, 16, 16), (8.0, 8.0, 8.0)), 6,
                                  dict(omega=(1.0, 1.0, 1.0), c0=5.0, ddi=True, cdd=0.5,
                                       c_extra=ce)),
              spin2_2d_fp32=(((32, 16), (10.0, 8.0)), 2,
                             dict(omega=(1.0, 1.0), c0=5.0, c1=-0.3, precision="fp32")))[case]
    (n, L), F, extra = kw
    og, _ = grids(n, L)
    psi = random_psi(og.shape, 2 * F + 1, seed=13, scale=0.2)
    ws = []
    for _ in range(2):
        _, g = pair(n, L, F, psi, 0.002, False, **extra)
        if case.endswith("split"):
            from becgpu import _lib as Lb
            Lb.lib().bec_set_xsite(g.handle, 0)
        ws.append(g)
    a, b = ws
    for core in (A.strang_core, A.yoshida_core):
        start = a.get_psi()
        a.copy_slot(2, 0)
        core(a, 0.002)
        b.swap_slots(2, 0)
        core(b, 0.002, src_slot=2)
        assert np.array_equal(a.get_psi(), b.get_psi())
        b.swap_slots(0, 2)                      # the source slot kept the starting state
        assert np.array_equal(b.get_psi(), start)
        b.swap_slots(0, 2)
        assert np.array_equal(b.get_psi(), a.get_psi())
    for w in ws:
        w.close()
