"""compute_flux (reference src/problem.jl:458-487) and pl_interpolate (src/utils.jl:23-27)
answered on the device, for one query or a batch of them.

Both take node indices 0-based.  A scalar query returns a float (a tuple over the
variables for an FVMSystem, as the reference); arrays of queries return arrays
(shape (n,) or, for a system, (n, nv)).  u may be a host array (an (nv, n_points)
matrix for a system) or a DeviceArray already in the device layout.
"""
import numpy as np

from . import _lib, forms
from .problem import SteadyFVMProblem
from .system import FVMSystem


def _device_state(prob, u):
    if isinstance(u, _lib.DeviceArray):
        return u
    if isinstance(prob, FVMSystem):
        return _lib.DeviceArray.from_host(prob.flat(u))
    return _lib.DeviceArray.from_host(np.asarray(u, np.float64))


def compute_flux(prob, i, j, u, t):
    """q(x, y, t, alpha, beta, gamma, p) . n across the edge (i, j) (problem.jl:458-487):
    (x, y) the edge midpoint, n the clockwise normal of i -> j, the shape function
    from the triangle on the right of i -> j, or on its left for a boundary edge
    whose domain lies on the left."""
    if isinstance(prob, SteadyFVMProblem):
        prob = prob.problem
    scalar = np.ndim(i) == 0 and np.ndim(j) == 0
    qi = np.ascontiguousarray(np.atleast_1d(i), np.int32)
    qj = np.ascontiguousarray(np.atleast_1d(j), np.int32)
    if qi.shape != qj.shape:
        raise ValueError("compute_flux: i and j must have the same shape")
    n = len(qi)
    ctx = prob.device_context()
    d_u = _device_state(prob, u)
    d_i = _lib.DeviceArray.from_host(qi)
    d_j = _lib.DeviceArray.from_host(qj)
    L = _lib.hip_lib()
    if isinstance(prob, FVMSystem):
        nv = prob.n_vars
        out = _lib.DeviceArray(max(n * nv, 1))
        _lib.check(L.fvm_system_compute_flux(ctx.handle, n, d_i.ptr, d_j.ptr, d_u.ptr, float(t),
                                             out.ptr, None), "compute_flux")
        r = out.to_host()[:n * nv].reshape(n, nv)
        return tuple(float(v) for v in r[0]) if scalar else r
    dq = None
    if ctx.flux_kind in (_lib.FLUX_D_EDGE, _lib.FLUX_LINEAR):
        tri = prob.mesh.triangulation
        ok = (qi >= 0) & (qi < tri.num_points()) & (qj >= 0) & (qj < tri.num_points())
        ii, jj = np.where(ok, qi, 0), np.where(ok, qj, 0)  # unknown nodes are the kernel's error
        mx = (tri.x[ii] + tri.x[jj]) / 2
        my = (tri.y[ii] + tri.y[jj]) / 2
        if ctx.flux_kind == _lib.FLUX_D_EDGE:
            dq = _lib.DeviceArray.from_host(forms.eval_xy(ctx.physics["d_xy"], mx, my))
        else:
            dq = _lib.DeviceArray.from_host(np.ascontiguousarray(ctx.physics["a_xy"](mx, my).T.ravel()))
    out = _lib.DeviceArray(max(n, 1))
    _lib.check(L.fvm_compute_flux(ctx.handle, n, d_i.ptr, d_j.ptr, d_u.ptr, float(t),
                                  dq.ptr if dq is not None else None, out.ptr, None), "compute_flux")
    r = out.to_host()[:n]
    return float(r[0]) if scalar else r


def pl_interpolate(prob, T, u, x, y):
    """alpha x + beta y + gamma of u on the mesh triangle T = (i, j, k) (any rotation
    of its counter-clockwise order) at (x, y) (utils.jl:23-27).  T may be an (n, 3)
    array with x, y of length n."""
    if isinstance(prob, SteadyFVMProblem):
        prob = prob.problem
    Ts = np.ascontiguousarray(np.asarray(T, np.int64).reshape(-1, 3), np.int32)
    scalar = np.ndim(x) == 0 and np.asarray(T).ndim == 1
    px = np.ascontiguousarray(np.atleast_1d(x), np.float64)
    py = np.ascontiguousarray(np.atleast_1d(y), np.float64)
    n = len(Ts)
    if px.shape != (n,) or py.shape != (n,):
        raise ValueError("pl_interpolate: one (x, y) per triangle")
    nv = prob.n_vars if isinstance(prob, FVMSystem) else 1
    d_u = _device_state(prob, u)
    d_t = _lib.DeviceArray.from_host(Ts.ravel())
    d_x = _lib.DeviceArray.from_host(px)
    d_y = _lib.DeviceArray.from_host(py)
    out = _lib.DeviceArray(max(n * nv, 1))
    _lib.check(_lib.hip_lib().fvm_pl_interpolate(prob.mesh.handle, nv, n, d_t.ptr, d_x.ptr, d_y.ptr,
                                                 d_u.ptr, out.ptr, None), "pl_interpolate")
    r = out.to_host()[:n * nv]
    if nv == 1:
        return float(r[0]) if scalar else r
    r = r.reshape(n, nv)
    return r[0].copy() if scalar else r
