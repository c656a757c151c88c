"""bec_spin_density (_compute_spin_density!, observables.jl:61-88) against the oracle.

The oracle's spin density is pinned separately (tests/test_oracle_pins.py: equal to the
expectation values of the spin matrices).  FP64 workspaces agree to round-off (1e-14
relative to the largest component); FP32-bulk workspaces hold psi in FP32 and form the
density in FP64, so it equals the oracle's density of the FP32-rounded spinor to round-off.
"""
import math

import numpy as np
import pytest

from oracle import spinor_oracle as O
from helpers import pair, random_psi, grids

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("F", [1, 2, 6])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("n,L", [((32,), (10.0,)), ((16, 24), (8.0, 9.0)), ((8, 16, 8), (6.0, 8.0, 6.0))])
def test_spin_density_matches_oracle(F, precision, n, L):
    D = 2 * F + 1
    og, _ = grids(n, L)
    psi = random_psi(og.shape, D, 7 + F) * O.gaussian(og, tuple(l / 6 for l in L))[None]
    psi /= math.sqrt(np.sum(np.abs(psi) ** 2) * og.dV)
    _, gws = pair(n, L, F, psi, 1e-3, False, precision=precision)
    got = gws.spin_density()
    ref_psi = psi if precision == "fp64" else psi.astype(np.complex64).astype(complex)
    ref = O.spin_density(ref_psi, F)
    tol = 1e-13
    scale = max(float(np.max(np.abs(r))) for r in ref)
    for g, r in zip(got, ref):
        assert g.shape == r.shape
        assert float(np.max(np.abs(g - r))) <= tol * scale
