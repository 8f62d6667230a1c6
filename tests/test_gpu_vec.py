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
