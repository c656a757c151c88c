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


def test_yoshida_driver_swap_equals_copy():
    """run_simulation_yoshida moves psi_strang with a slot swap when every instance is
    active (a copy otherwise): the same bits as the copy."""
    og, _ = grids((32, 32), (10.0, 10.0))
    psi = random_psi(og.shape, 3, seed=9, scale=0.2)
    outs = []
    for use_copy in (False, True):
        _, gws = pair((32, 32), (10.0, 10.0), 1, psi, 0.002, False, omega=(1.0, 1.0), c0=5.0,
                      c1=-0.3)
        if use_copy:
            gws.swap_slots = lambda a, b: gws.copy_slot(a, b)
        r = A.run_simulation_yoshida(gws, t_end=0.05, save_interval=0.025,
                                     keep_snapshots=False)
        outs.append((gws.get_psi(), r["n_accepted"], r["n_rejected"]))
        gws.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and outs[0][1:] == outs[1][1:]
