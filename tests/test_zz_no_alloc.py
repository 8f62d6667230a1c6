"""fvm_vec_mdot_noalloc, the entry point that allocates nothing: its k dot products
land in the caller's device array and its partial sums in the library's static
per-device scratch.  It must equal fvm_vec_mdot bit for bit (the same kernels, grid
and summation order), on two calls in a row, the second with a smaller input, and
reject a k beyond what the scratch holds."""
import ctypes

import numpy as np
import pytest

from finitevolumemethod import _lib

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def _noalloc(xs, y, out):
    k = len(xs)
    ptrs = (_vp * k)(*[x.ptr for x in xs])
    rc = _lib.hip_lib().fvm_vec_mdot_noalloc(y.n, k, ptrs, y.ptr, out.ptr, None)
    _lib.synchronize()
    return rc


def test_noalloc_mdot_equals_mdot_on_two_calls(gpu):
    rng = np.random.default_rng(5)
    out = _lib.DeviceArray(64)
    for n, k in ((300_001, 64), (1_001, 5)):  # the second call smaller in n and k
        X = [_lib.DeviceArray.from_host(rng.standard_normal(n)) for _ in range(k)]
        y = _lib.DeviceArray.from_host(rng.standard_normal(n))
        ref = _lib.mdot(X, y)
        assert _noalloc(X, y, out) == 0, _lib.last_error()
        got = out.to_host()[:k]
        assert np.array_equal(got, ref), (n, k)
        Xh = np.array([x.to_host() for x in X])
        assert np.all(np.abs(got - Xh @ y.to_host()) <= 1e-14 * (np.abs(Xh) @ np.abs(y.to_host())))


def test_noalloc_mdot_rejects_more_vectors_than_the_scratch_holds(gpu):
    y = _lib.DeviceArray.from_host(np.ones(10))
    X = [y] * 65
    out = _lib.DeviceArray(65)
    assert _noalloc(X, y, out) != 0
    assert "1..64" in _lib.last_error()
