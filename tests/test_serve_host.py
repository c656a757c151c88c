"""Host side of the request server: the seeded request stream (no device calls)."""
import numpy as np

from becgpu.serve import SIDES12, SIDES3, Request, _pool_key, poisson_mix


def test_poisson_mix_is_seeded_and_shaped():
    a = poisson_mix(4000, rate=200.0, seed=42)
    b = poisson_mix(4000, rate=200.0, seed=42)
    assert [r.__dict__ for r in a] == [r.__dict__ for r in b]
    assert poisson_mix(50, seed=43)[0].__dict__ != a[0].__dict__
    t = np.array([r.arrival for r in a])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    # exponential gaps of mean 1/200 s: mean within 5 standard errors
    assert abs(gaps.mean() - 1 / 200) < 5 * (1 / 200) / np.sqrt(len(gaps))
    nd = np.array([len(r.n) for r in a])
    F = np.array([r.F for r in a])
    for arr, vals, probs in ((nd, (1, 2, 3), (0.5, 0.35, 0.15)), (F, (1, 2, 6), (0.6, 0.25, 0.15))):
        for v, p in zip(vals, probs):
            frac = np.mean(arr == v)
            assert abs(frac - p) < 5 * np.sqrt(p * (1 - p) / len(a)), (v, frac, p)
    for r in a:
        assert r.kind in ("itp", "rt")
        assert 500 <= r.n_steps <= 5000
        assert 5.0 <= r.c0 <= 20.0 and -1.0 <= r.c1 <= 1.0
        assert len(set(r.n)) == 1
        assert r.n[0] in (SIDES3 if len(r.n) == 3 else SIDES12)
        assert r.dt == (0.005 if r.kind == "itp" else 0.001)
    kinds = np.array([r.kind == "itp" for r in a])
    assert abs(kinds.mean() - 0.5) < 5 * np.sqrt(0.25 / len(a))


def test_pool_key_separates_shapes_and_modes():
    base = dict(rid=0, kind="itp", n=(64,), L=(20.0,), F=1, c0=10.0, c1=0.0, n_steps=10, dt=0.005)
    k0 = _pool_key(Request(**base))
    assert _pool_key(Request(**{**base, "c0": 3.0, "rid": 5, "n_steps": 99})) == k0
    for ch in ({"kind": "rt"}, {"n": (128,)}, {"F": 2}, {"precision": "fp32"}, {"L": (10.0,)}):
        assert _pool_key(Request(**{**base, **ch})) != k0
