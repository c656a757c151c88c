"""bec_fft_gradient_noalloc (the entry point that allocates no device memory) against
bec_fft_gradient, through the C ABI.

Both run the same kernels on the same data, so the gradients must agree bit for bit.  Two
calls in a row on the one shared buffer, the second on a smaller grid (a stale tail of the
first call's data must not leak into it), then the first shape again; every result is also
checked against the oracle's spectral gradient (fft_utils.jl:25-44) to 1e-12.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import spinor_oracle as O
import becgpu as G
from becgpu import _lib as L
from becgpu import api as A

pytestmark = pytest.mark.gpu


def _both(n, L_box, seed):
    gg = G.make_grid(G.GridConfig(tuple(n), tuple(L_box)))
    og = O.make_grid(n, L_box)
    f = np.ascontiguousarray(np.random.default_rng(seed).normal(size=gg.shape))
    cfg, keep = A._grid_config(gg)
    nd = gg.ndim
    ref = np.empty((nd,) + gg.shape)
    out = np.full((nd,) + gg.shape, np.nan)
    L.check(L.lib().bec_fft_gradient(C.byref(cfg), f.ctypes.data, ref.ctypes.data),
            "bec_fft_gradient")
    L.check(L.lib().bec_fft_gradient_noalloc(C.byref(cfg), f.ctypes.data, out.ctypes.data),
            "bec_fft_gradient_noalloc")
    oracle = O.fft_gradient(f, og)
    return out, ref, oracle


@pytest.mark.parametrize("shapes", [
    [((64, 32, 16), (8.0, 6.0, 5.0)), ((16, 8, 8), (4.0, 3.0, 3.0)), ((64, 32, 16), (8.0, 6.0, 5.0))],
    [((256, 128), (10.0, 6.0)), ((24, 40), (3.0, 5.0))],
    [((512,), (20.0,)), ((48,), (7.0,))],
])
def test_noalloc_gradient_equals_allocating_path(shapes):
    for k, (n, L_box) in enumerate(shapes):
        out, ref, oracle = _both(n, L_box, 10 + k)
        assert np.array_equal(out, ref), (n, float(np.max(np.abs(out - ref))))
        scale = max(1.0, float(np.max(np.abs(ref))))
        for d in range(len(n)):
            assert float(np.max(np.abs(out[d] - oracle[d]))) <= 1e-12 * scale


def test_noalloc_gradient_rejects_oversize_grid():
    gg = G.make_grid(G.GridConfig((128, 64, 64), (8.0, 6.0, 5.0)))   # 524288 sites > 2^18
    cfg, keep = A._grid_config(gg)
    f = np.zeros(gg.shape)
    out = np.zeros((3,) + gg.shape)
    rc = L.lib().bec_fft_gradient_noalloc(C.byref(cfg), f.ctypes.data, out.ctypes.data)
    assert rc == L.BEC_EINVAL
    assert b"exceeds the fixed buffer" in L.lib().bec_last_error()
