"""compute_flux (reference src/problem.jl:458-487) and pl_interpolate (src/utils.jl:23-27)
answered on the device, for one query or a batch of them.

Both take node indices 0-based.  A scalar query returns a float (a tuple over the
variables for an FVMSystem, as the reference); arrays of queries return arrays
(shape (n,) or, for a system, (n, nv)).  u may be a host array (an (nv, n_points)
matrix for a system) or a DeviceArray already in the device layout.
"""
import numpy as np

from . import _lib, forms
from .geometry import FVMGeometry
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


def _start_grid(mesh):
    """A G x G grid over the mesh's bounding box with, per cell, the mesh vertex nearest
    to the cell centre (built once per mesh): the start vertices of point location."""
    g = getattr(mesh, "_locate_grid", None)
    if g is not None:
        return g
    from scipy.spatial import cKDTree

    tri = mesh.triangulation
    verts = np.unique(tri.triangles)
    px, py = tri.x[verts], tri.y[verts]
    x0, x1, y0, y1 = px.min(), px.max(), py.min(), py.max()
    G = int(max(1, min(2048, np.sqrt(len(verts) / 4))))
    hx = (x1 - x0) / G if x1 > x0 else 1.0
    hy = (y1 - y0) / G if y1 > y0 else 1.0
    cx = x0 + (np.arange(G) + 0.5) * hx
    cy = y0 + (np.arange(G) + 0.5) * hy
    CX, CY = np.meshgrid(cx, cy, indexing="ij")
    _, nearest = cKDTree(np.stack([px, py], 1)).query(np.stack([CX.ravel(), CY.ravel()], 1))
    g = (x0, y0, hx, hy, G, verts[nearest].astype(np.int32))
    mesh._locate_grid = g
    return g


def find_triangle(obj, x, y, start=None):
    """The mesh triangle (i, j, k), counter-clockwise, containing each point (x, y), or
    (-1, -1, -1) outside the mesh: what the reference's tutorials obtain from
    DelaunayTriangulation's jump_and_march / find_triangle before calling
    pl_interpolate.  obj is an FVMGeometry or a problem; start is an initial vertex per
    point (or one for all), by default the vertex nearest to the point's cell of a grid
    over the mesh.  The search runs on the device (fvm_locate_points)."""
    mesh = obj if isinstance(obj, FVMGeometry) else obj.mesh
    tri = mesh.triangulation
    scalar = np.ndim(x) == 0 and np.ndim(y) == 0
    xs = np.ascontiguousarray(np.atleast_1d(x), np.float64)
    ys = np.ascontiguousarray(np.atleast_1d(y), np.float64)
    if xs.shape != ys.shape:
        raise ValueError("find_triangle: x and y must have the same shape")
    n = len(xs)
    if start is None:
        x0, y0, hx, hy, G, reps = _start_grid(mesh)
        ix = np.clip(((xs - x0) / hx).astype(np.int64), 0, G - 1)
        iy = np.clip(((ys - y0) / hy).astype(np.int64), 0, G - 1)
        st = reps[ix * G + iy]
    else:
        st = np.broadcast_to(np.asarray(start, np.int32), (n,))
    st = np.ascontiguousarray(st, np.int32)
    d_x, d_y, d_s = (_lib.DeviceArray.from_host(a) for a in (xs, ys, st))
    d_o = _lib.DeviceArray(max(n, 1), np.int32)
    _lib.check(_lib.hip_lib().fvm_locate_points(mesh.handle, n, d_x.ptr, d_y.ptr, d_s.ptr, d_o.ptr, None),
               "fvm_locate_points")
    t = d_o.to_host()[:n]
    out = np.full((n, 3), -1, np.int64)
    ok = t >= 0
    out[ok] = tri.triangles[t[ok]]
    return tuple(int(v) for v in out[0]) if scalar else out
