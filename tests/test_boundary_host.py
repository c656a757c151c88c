"""CPU tests of the C-ABI boundary (symbols, validation paths that run before any
device call) and of the host-side setup code of the becgpu package."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest
from sympy.physics.wigner import clebsch_gordan as s_cg, wigner_6j as s_6j

import becgpu as G
from becgpu import _lib as L
from becgpu import physics as P
from oracle import spinor_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "becgpu.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bec_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 20
    lib = C.CDLL(L.LIB_PATH)
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.SIGNATURES), set(syms) ^ set(L.SIGNATURES)


def test_line_plan_choice():
    """bec_line_plan (host only): built radix plans, runtime mixed-radix passes, and
    Bluestein's algorithm from a cost ratio of 2.5 (sides 214, 218, 226, 254, ...)."""
    lib = L.lib()
    for n in (8, 24, 40, 56, 128, 224, 512):
        assert lib.bec_line_plan(n) == 0, n
    for n in (6, 10, 22, 62, 72, 100, 118, 122, 188, 202, 1024):
        assert lib.bec_line_plan(n) == 1, n
    for n in (214, 218, 226, 254, 314, 452, 502, 508):
        assert lib.bec_line_plan(n) == 2, n
    assert lib.bec_line_plan(1) == L.BEC_EINVAL


def _cfg(**kw):
    c = L.BecConfig()
    c.ndim = kw.get("ndim", 1)
    n = kw.get("n", (64, 1, 1))
    for d in range(3):
        c.n[d] = n[d]
        c.L[d] = 10.0
    c.F = kw.get("F", 1)
    c.n_instances = kw.get("B", 1)
    c.precision = 0
    c.imaginary_time = 1
    c.normalize_every = 1
    c.enable_ddi = kw.get("ddi", 0)
    buf = (C.c_double * 4096)()
    c.x_tab = C.cast(buf, C.POINTER(C.c_double))
    c.k_tab = C.cast(buf, C.POINTER(C.c_double))
    return c, buf


@pytest.mark.parametrize("kw,code", [
    (dict(n=(63, 1, 1)), L.BEC_EINVAL),
    (dict(ndim=4), L.BEC_EINVAL),
    (dict(F=7), L.BEC_EUNSUP),                  # more than 13 components: not built
    (dict(F=-1), L.BEC_EINVAL),                 # SpinSystem(F) needs F >= 0 (types.jl:43)
    # any even side is built; what one block's LDS cannot hold is refused: 13 whole x rows
    # of 2048 points, a 16384-point line, a y line of 16384 points
    (dict(F=6, n=(2048, 1, 1)), L.BEC_EUNSUP),
    (dict(n=(16384, 1, 1)), L.BEC_EUNSUP),
    (dict(ndim=2, n=(8, 16384, 1)), L.BEC_EUNSUP),
    (dict(ndim=3, n=(1024, 1024, 1024)), L.BEC_EUNSUP),  # 2^31 values per instance
    (dict(B=0), L.BEC_EINVAL),
])
def test_create_rejects_invalid_configs(kw, code):
    c, buf = _cfg(**kw)
    h = C.c_void_p()
    rc = L.lib().bec_ws_create(C.byref(c), C.byref(h))
    assert rc == code
    assert h.value is None
    assert L.lib().bec_last_error().decode()


def test_workspace_bytes_scales_with_instances():
    c, buf = _cfg(ndim=3, n=(64, 64, 64), F=1, B=1)
    b1 = L.lib().bec_workspace_bytes(C.byref(c))
    c.n_instances = 4
    b4 = L.lib().bec_workspace_bytes(C.byref(c))
    assert b1 >= 2 * 3 * 64 ** 3 * 16
    assert b4 > 3.9 * b1


# --- host setup (physics.py) vs independent implementations ------------------------
def test_host_grid_matches_oracle():
    gc = G.GridConfig((16, 8, 4), (3.0, 5.0, 7.0))
    g = G.make_grid(gc)
    og = O.make_grid((16, 8, 4), (3.0, 5.0, 7.0))
    for d in range(3):
        assert np.allclose(g.x[d], og.x[d], rtol=0, atol=1e-15)
        assert np.allclose(g.k[d], og.k[d], rtol=1e-15, atol=1e-15)
    assert np.allclose(g.k_squared, og.k2, rtol=1e-14)
    with pytest.raises(ValueError):
        G.GridConfig((5,), (1.0,))


@pytest.mark.parametrize("F", [1, 2, 6])
def test_host_racah_cg_matches_sympy(F):
    t = P.precompute_cg_table(F)
    for (l, M, m1, m2), v in list(t.items())[::7]:
        assert v == pytest.approx(float(s_cg(F, F, l, m1, m2, M)), abs=1e-13)
    assert P.wigner_6j(F, F, 2, F, F, 0) == pytest.approx(float(s_6j(F, F, 2, F, F, 0)), abs=1e-13)
    if F == 6:
        # every exactly-nonzero coefficient is present (773 at F=6); the log-space
        # Racah sum may leave an exact zero at roundoff level (|v| ~ 1e-15)
        exact = {(l, M, m1, M - m1) for l in range(0, 13, 2) for M in range(-l, l + 1)
                 for m1 in range(-6, 7) if abs(M - m1) <= 6
                 and float(s_cg(6, 6, l, m1, M - m1, M)) != 0.0}
        assert len(exact) == 773 and exact <= set(t)
        assert all(abs(t[k]) < 1e-13 for k in set(t) - exact)


def test_host_init_states_match_oracle():
    gc = G.GridConfig((32, 16), (10.0, 8.0))
    g = G.make_grid(gc)
    og = O.make_grid((32, 16), (10.0, 8.0))
    for F in (1, 2):
        for st in ("polar", "ferromagnetic", "uniform", "antiferromagnetic"):
            a = G.init_psi(g, G.SpinSystem(F), st)
            b = O.init_psi(og, F, st)
            assert np.allclose(a, b, atol=1e-14)
    V = O.harmonic_potential(og, (1.0, 1.2))
    a = G.init_psi_thomas_fermi(g, G.SpinSystem(1), G.HarmonicTrap(1.0, 1.2), 10.0)
    b = O.thomas_fermi_psi(og, 1, V, 10.0)
    assert np.allclose(a, b, atol=1e-13)


def test_random_state_is_seeded_and_normalised():
    g = G.make_grid(G.GridConfig((16,), (5.0,)))
    a = G.init_psi(g, G.SpinSystem(1), "random", seed=43)
    b = G.init_psi(g, G.SpinSystem(1), "random", seed=43)
    c = G.init_psi(g, G.SpinSystem(1), "random", seed=44)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.sum(np.abs(a) ** 2) * G.cell_volume(g) == pytest.approx(1.0, rel=1e-12)
    z = P.SplitMix64(1).randn_complex(200000)
    assert np.mean(np.abs(z) ** 2) == pytest.approx(1.0, abs=0.01)


def test_itp_validation_errors():
    from becgpu import api as A
    with pytest.raises(ValueError):
        A._validate_itp_zeeman(G.ZeemanParams(1e5, 0.0), 6, 0.01)
    with pytest.raises(ValueError):
        A._validate_itp_interactions(G.InteractionParams(1e6, 0.0), 1, 0.01)
    A._validate_itp_interactions(G.InteractionParams(10.0, 0.5), 1, 0.01)


def test_classify_phase_host():
    g = G.make_grid(G.GridConfig((16,), (8.0,)))
    ferro = G.init_psi(g, G.SpinSystem(1), "ferromagnetic")
    polar = G.init_psi(g, G.SpinSystem(1), "polar")
    assert P.classify_phase(ferro, 1, g)["phase"] == "ferromagnetic"
    assert P.classify_phase(polar, 1, g)["phase"] == "polar"


@pytest.mark.parametrize("F", [1, 2, 6])
def test_host_loss_rates_match_oracle(F):
    """Host dipolar relaxation rates (log-factorial Racah CG) vs the oracle's sympy CG."""
    from becgpu import physics as P
    assert np.allclose(P.dipolar_relaxation_rates(F, 0.7), O.dipolar_relaxation_rates(F, 0.7),
                       rtol=1e-12, atol=1e-15)


def test_adaptive_dt_params_validation():
    """types.jl:343-361 / test_adaptive_dt.jl:28-47."""
    p = G.AdaptiveDtParams()
    assert (p.dt_init, p.dt_min, p.dt_max, p.tol, p.error_mode) == (0.001, 1e-5, 0.01, 1e-3, "step_change")
    assert G.AdaptiveDtParams(error_mode="richardson").error_mode == "richardson"
    for bad in (dict(dt_init=-1.0), dict(dt_min=0.1, dt_max=0.01), dict(tol=-1.0), dict(error_mode="foo")):
        with pytest.raises(ValueError):
            G.AdaptiveDtParams(**bad)
