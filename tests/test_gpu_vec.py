"""Device vector kernels used by the integrators and Krylov solvers: the fused
multi-dot and multi-axpy against NumPy (1e-14 relative), their bitwise
reproducibility, and GMRES(m) with classical Gram-Schmidt applied twice on a
nonsymmetric operator against a dense solve."""
import numpy as np
import pytest

from finitevolumemethod import _lib
from finitevolumemethod.linsolve import gmres_device

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,k", [(1, 1), (1000, 5), (123457, 16), (300001, 17), (50000, 64)])
def test_mdot_and_maxpy_match_numpy(gpu, n, k):
    rng = np.random.default_rng(n + k)
    X = rng.standard_normal((k, n))
    y = rng.standard_normal(n)
    dX = [_lib.DeviceArray.from_host(x) for x in X]
    dy = _lib.DeviceArray.from_host(y)
    got = _lib.mdot(dX, dy)
    ref = X @ y
    scale = np.abs(X) @ np.abs(y)
    assert np.all(np.abs(got - ref) <= 1e-14 * scale + 1e-300)
    assert np.array_equal(_lib.mdot(dX, dy), got)  # reproducible
    c = rng.standard_normal(k)
    _lib.maxpy(dy, c, dX)
    ref2 = y + c @ X
    assert np.all(np.abs(dy.to_host() - ref2) <= 1e-14 * (np.abs(y) + np.abs(c) @ np.abs(X)) + 1e-300)


def test_gmres_solves_a_nonsymmetric_system(gpu):
    n = 400
    rng = np.random.default_rng(3)
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    dA_cols = None  # apply A through host for this small dense check

    def apply_A(v, out):
        out.upload(A @ v.to_host())

    x = _lib.DeviceArray(n)
    it = gmres_device(apply_A, _lib.DeviceArray.from_host(b), x, rtol=1e-12, restart=30, maxiter=300)
    assert it <= 300
    xs = np.linalg.solve(A, b)
    assert np.linalg.norm(x.to_host() - xs) <= 1e-9 * np.linalg.norm(xs)
    del dA_cols


def test_reductions_from_several_host_threads(gpu):
    """Dot products and error norms issued concurrently from four host threads,
    each on its own stream, equal the same reductions done one at a time (the
    reductions share a scratch area, which calls hold for their whole duration)."""
    import ctypes
    import threading

    from finitevolumemethod import _lib

    rng = np.random.default_rng(11)
    n = 300_000
    xs = [rng.standard_normal(n) for _ in range(4)]
    d = [_lib.DeviceArray.from_host(x) for x in xs]
    ref = [_lib.dot(a, a) for a in d]
    refn = [_lib.err_norm(a, a, a, 1e-6, 1e-3) for a in d]
    out = [[None, None] for _ in d]
    errs = []

    L = _lib.hip_lib()
    streams = []
    for _ in range(4):
        h = ctypes.c_void_p()
        _lib.check(L.fvm_stream_create(ctypes.byref(h)), "fvm_stream_create")
        streams.append(h.value)

    def work(k):
        try:
            for _ in range(20):
                out[k][0] = _lib.dot(d[k], d[k], stream=streams[k])
                out[k][1] = _lib.err_norm(d[k], d[k], d[k], 1e-6, 1e-3, stream=streams[k])
        except Exception as e:  # noqa: BLE001
            errs.append(e)

    th = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for h in streams:
        _lib.check(L.fvm_stream_destroy(h), "fvm_stream_destroy")
    for k in range(4):
        assert out[k][0] == ref[k] and out[k][1] == refn[k]


class _View:
    """n doubles of a device buffer starting at element off (8-byte aligned only when
    off is odd), for the unaligned path of the vector kernels."""

    def __init__(self, arr, off, n):
        self.arr, self.n, self.ptr = arr, n, arr.ptr + 8 * off

    def to_host(self):
        return self.arr.to_host()[(self.ptr - self.arr.ptr) // 8:][: self.n]


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 123457, 300002])
@pytest.mark.parametrize("off", [0, 1], ids=["aligned", "offset"])
def test_lincomb_equals_left_to_right_sum(gpu, n, off):
    """Every term count, 16-byte aligned and odd-offset vectors, even and odd n:
    out = c0 x0 + c1 x1 + ... summed left to right, which NumPy's elementwise
    operations reproduce exactly (separately rounded products and sums)."""
    rng = np.random.default_rng(n + 7 * off)
    for nt in range(1, 9):
        X = rng.standard_normal((nt, n))
        c = rng.standard_normal(nt)
        bufs = [_lib.DeviceArray.from_host(np.concatenate([np.zeros(off), x, [0.0]])) for x in X]
        xs = [_View(b, off, n) for b in bufs]
        out_buf = _lib.DeviceArray(n + off + 1)
        out = _View(out_buf, off, n)
        _lib.lincomb(out, c, xs)
        ref = c[0] * X[0]
        for k in range(1, nt):
            ref = ref + c[k] * X[k]
        assert np.array_equal(out.to_host(), ref), nt
        # out aliasing the first input, as the integrators use it
        _lib.lincomb(xs[0], c, xs)
        assert np.array_equal(xs[0].to_host(), ref), nt


@pytest.mark.parametrize("n", [1, 1001, 300001])
def test_fused_error_norm_equals_lincomb_then_norm(gpu, n):
    rng = np.random.default_rng(n)
    k = [_lib.DeviceArray.from_host(rng.standard_normal(n)) for _ in range(7)]
    u0 = _lib.DeviceArray.from_host(rng.standard_normal(n))
    u1 = _lib.DeviceArray.from_host(rng.standard_normal(n))
    c = rng.standard_normal(7) * 1e-3
    err = _lib.DeviceArray(n)
    _lib.lincomb(err, c, k)
    ref = _lib.err_norm(err, u0, u1, 1e-6, 1e-3)
    got = _lib.lincomb_err_norm(c, k, u0, u1, 1e-6, 1e-3)
    assert got == ref
    e = err.to_host()
    sc = 1e-6 + 1e-3 * np.maximum(np.abs(u0.to_host()), np.abs(u1.to_host()))
    assert abs(got - np.sqrt(np.mean((e / sc) ** 2))) <= 1e-13 * got
