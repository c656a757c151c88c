"""1-D split steps repeated inside one launch (bec_split_step on a 1-D grid whose step is a
single x pass): n steps in one call must give the bits of n one-step calls, and the oracle's
result at the FP64 bars.

The reference runs split_step! once per step (split_step.jl:10-35; find_ground_state's loop,
ground_state.jl:80-163); the repeated launch keeps the line in LDS between steps, so these
tests compare it with the per-step launches bit for bit and with the oracle at 1e-10 / 1e-12.
"""
import numpy as np
import pytest

from oracle import spinor_oracle as O
from becgpu import api as A
from helpers import pair, random_psi, rel_l2

pytestmark = pytest.mark.gpu


def _run(gws, n_steps, chunks):
    """Advance n_steps split steps in calls of the given sizes."""
    done = 0
    for c in chunks:
        A.split_step(gws, c)
        done += c
    assert done == n_steps
    return gws.get_psi()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("itp", [True, False])
@pytest.mark.parametrize("n", [16, 64, 96, 256])
def test_repeated_steps_equal_single_steps(n, itp, precision):
    """F = 1 with trap, Zeeman, contact and spin mixing: 37 steps in one call, in calls of
    1, 5, 31 steps, and one step at a time give the same psi bit for bit."""
    L = 20.0
    psi = random_psi((n,), 3, seed=11 + n, scale=0.3)
    kw = dict(omega=(1.0,), p=0.05, q=0.1, c0=10.0, c1=-0.5, precision=precision)
    outs = []
    for chunks in ([37], [1, 5, 31], [1] * 37):
        _, gws = pair((n,), (L,), 1, psi, 0.005, itp, **kw)
        outs.append(_run(gws, 37, chunks))
        gws.close()
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("itp", [True, False])
def test_repeated_steps_loss_and_scalar_gas(itp):
    """Dipolar-relaxation and three-body loss in real time (ops after the kinetic step) and
    the scalar gas F = 0."""
    psi = random_psi((64,), 3, seed=5, scale=0.3)
    outs = []
    for chunks in ([40], [1] * 40):
        _, gws = pair((64,), (20.0,), 1, psi, 0.002, itp, omega=(1.0,), c0=10.0, c1=0.3,
                      loss=None if itp else (0.05, 0.01))
        outs.append(_run(gws, 40, chunks))
        gws.close()
    assert np.array_equal(outs[0], outs[1])
    psi0 = random_psi((64,), 1, seed=6, scale=0.3)
    outs = []
    for chunks in ([25], [1] * 25):
        _, gws = pair((64,), (20.0,), 0, psi0, 0.002, itp, omega=(1.0,), c0=10.0)
        outs.append(_run(gws, 25, chunks))
        gws.close()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("itp", [True, False])
def test_repeated_steps_match_oracle(itp):
    """200 repeated steps against the oracle's split_step! loop: psi to 1e-10 (relative L2
    over all components; in imaginary time one component decays to 1e-10 of the others, so
    its own relative error measures only rounding) and the total energy to 1e-12."""
    n, L = 64, 20.0
    psi = random_psi((n,), 3, seed=21, scale=0.3)
    ows, gws = pair((n,), (L,), 1, psi, 0.005, itp, omega=(1.0,), q=0.1, c0=10.0, c1=-0.5)
    A.split_step(gws, 200)
    for _ in range(200):
        O.split_step(ows)
    assert rel_l2(gws.get_psi(), ows.psi) < 1e-10
    e_gpu = A.total_energy(gws)
    e_ref = O.total_energy(ows)
    assert abs(e_gpu - e_ref) <= 1e-12 * abs(e_ref)
    gws.close()
