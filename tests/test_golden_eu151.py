"""End-to-end 3-D ITP pinned by the reference's own golden output
(examples/eu151/density_64.json, fixture tests/golden/eu151_density_64.json):
151Eu F=6, 64^3, box 20, trap (1, 1, 130/110), c0 = 4 pi (a_s/a_ho) 5e4, c1 = 0,
p = 100, dt = 0.005, n_steps 20000, tol 1e-9, ferromagnetic start, no DDI
(examples/eu151/eu151_setup.jl:4-24, eu151_params.jl:5-18)."""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "eu151_density_64.json")))

HBAR, AMU, A0 = 1.054571817e-34, 1.66053906660e-27, 5.29177210903e-11


def eu_c0():
    m = 150.919857 * AMU
    a_ho = math.sqrt(HBAR / (m * 2 * math.pi * 110.0))
    return 4 * math.pi * (110 * A0 / a_ho) * 50000


def check_density(n):
    """n: total density, numpy layout (z, y, x)."""
    assert np.max(n) == pytest.approx(GOLD["n_peak"], rel=1e-9)
    ref_peak = max(GOLD["n_x"])
    for key, cut in (("n_x", n[32, 32, :]), ("n_y", n[32, :, 32]), ("n_z", n[:, 32, 32])):
        assert np.max(np.abs(cut - np.array(GOLD[key]))) <= 1e-9 * ref_peak, key
    sl = np.array(GOLD["slice_xy"]).reshape(64, 64)   # (y, x) at z = 32
    assert np.max(np.abs(n[32] - sl)) <= 1e-9 * ref_peak


@pytest.mark.gpu
def test_gpu_eu151_ground_state_matches_golden():
    import becgpu as G
    grid = G.make_grid(G.GridConfig((64, 64, 64), (20.0, 20.0, 20.0)))
    r = G.find_ground_state(grid=grid, atom=G.Eu151, interactions=G.InteractionParams(eu_c0(), 0.0),
                            zeeman=G.ZeemanParams(100.0, 0.0),
                            potential=G.HarmonicTrap(1.0, 1.0, 130.0 / 110.0), dt=0.005,
                            n_steps=20000, tol=1e-9, initial_state="ferromagnetic")
    assert r["converged"]
    psi = r["workspace"].get_psi()
    pops = np.sum(np.abs(psi.reshape(13, -1)) ** 2, axis=1) * (20.0 / 64) ** 3
    assert np.allclose(pops, GOLD["comp_pops"], atol=1e-12)
    check_density(np.sum(np.abs(psi) ** 2, axis=0))


@pytest.mark.skipif(os.environ.get("BEC_SLOW") != "1", reason="oracle run takes ~2.5 min; BEC_SLOW=1")
def test_oracle_eu151_ground_state_matches_golden():
    """With c1 = 0, no DDI and a ferromagnetic start every m < 6 component stays exactly
    zero (the diagonal step multiplies zeros, spin mixing is skipped), so the scalar
    (F=0) run is the same arithmetic on the m=+6 component."""
    from oracle import spinor_oracle as O
    g = O.make_grid((64, 64, 64), (20.0, 20.0, 20.0))
    psi = O.init_psi(g, 0, "ferromagnetic")
    ws = O.make_workspace(g, 0, psi, 0.005, True, omega=(1, 1, 130 / 110), c0=eu_c0(),
                          n_steps=20000)
    r = O.find_ground_state(ws, 20000, 1e-9)
    assert r["converged"]
    check_density(np.abs(ws.psi[0]) ** 2)
