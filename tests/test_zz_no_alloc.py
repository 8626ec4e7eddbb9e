"""The proposal network's forward pass without allocation (GPU): upbp_mlp_forward_noalloc, whose
input staging and outputs live in a fixed per-GPU device array, must give exactly the
outputs of upbp_mlp_forward (the same kernels on the same inputs), on two calls in a row -- the
second smaller, so that it reads nothing the first left behind -- and reject a size beyond its
maximum."""
import ctypes as C

import numpy as np
import pytest

from gpuupbp_amd import capi
from gpuupbp_amd import renderer as gpu

pytestmark = pytest.mark.gpu


def _samples(rng, n):
    x = rng.uniform(0, 1, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.ascontiguousarray(np.hstack([x, d]), np.float32)


def _noalloc(m, s):
    out = np.full((len(s), 40), np.nan, np.float32)
    rc = capi.lib().upbp_mlp_forward_noalloc(m.h, s.ctypes.data_as(C.c_void_p), len(s),
                                              out.ctypes.data_as(C.c_void_p))
    assert rc == 0, rc
    return out


def test_noalloc_forward_equals_forward():
    lib = capi.lib()
    cap = lib.upbp_mlp_noalloc_max()
    assert cap == 32768
    m = gpu.Mlp(seed=11)
    rng = np.random.default_rng(9)
    big, small = _samples(rng, 20000), _samples(rng, 1234)  # 1234: not a multiple of 16
    for s in (big, small):
        want = m.forward(s)
        got = _noalloc(m, s)
        assert np.isfinite(got).all()
        assert np.array_equal(got, want)
    # the largest size it takes, then one beyond: a clear capacity error, outputs untouched
    s = _samples(rng, cap)
    assert np.array_equal(_noalloc(m, s), m.forward(s))
    over = _samples(rng, cap + 1)
    out = np.zeros((cap + 1, 40), np.float32)
    rc = lib.upbp_mlp_forward_noalloc(m.h, over.ctypes.data_as(C.c_void_p), cap + 1, out.ctypes.data_as(C.c_void_p))
    assert rc == -6 and not out.any()  # UPBP_E_CAPACITY
    m.close()
