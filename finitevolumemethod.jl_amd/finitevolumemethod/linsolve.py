"""Krylov solvers over device vectors (matrix-free operators)."""
import math

import numpy as np

from . import _lib


_WORKSPACES = {}


def _workspace(n, restart):
    """The Krylov basis and two work vectors, kept per (n, restart) between calls:
    implicit steppers call GMRES for every Newton iteration, and allocating (and
    freeing, which synchronises the device) restart + 3 vectors each time cost more
    than the iterations on small meshes.  Callers must not nest calls of one size."""
    ws = _WORKSPACES.get((n, restart))
    if ws is None:
        if len(_WORKSPACES) >= 8:
            _WORKSPACES.clear()
        ws = ([_lib.DeviceArray(n) for _ in range(restart + 1)], _lib.DeviceArray(n), _lib.DeviceArray(n))
        _WORKSPACES[(n, restart)] = ws
    return ws


def library_arnoldi(op_kind, handle, u, t, a, b):
    """The Arnoldi step of gmres_device for A = a I + b J(u, t), J the Jacobian of a
    library operator (fvm_arnoldi_step: the product and both Gram-Schmidt passes
    with one host synchronisation; the same kernels as the step it replaces).  With a
    right preconditioner diag(scale), gmres_device passes scale and a work vector and
    the step is fvm_arnoldi_step_scaled, A diag(scale) V[j]."""
    L = _lib.hip_lib()
    buf = np.empty(2 * 64 + 1)

    def step(V, j, w, scale=None, tmp=None):
        k = j + 1
        ptrs = (_lib._vp * k)(*[v.ptr for v in V[:k]])
        if scale is None:
            _lib.check(L.fvm_arnoldi_step(int(op_kind), handle, w.n, u.ptr, float(t), float(a), float(b), j, ptrs,
                                          w.ptr, _lib.ptr(buf), None), "fvm_arnoldi_step")
        else:
            _lib.check(L.fvm_arnoldi_step_scaled(int(op_kind), handle, w.n, u.ptr, float(t), float(a), float(b),
                                                 scale.ptr, tmp.ptr, j, ptrs, w.ptr, _lib.ptr(buf), None),
                       "fvm_arnoldi_step_scaled")
        return buf[:k].copy(), buf[k:2 * k].copy(), float(buf[2 * k])

    return step


def gmres_device(apply_A, b, x, rtol=1e-6, restart=50, maxiter=1000, arnoldi=None, precond=None):
    """Restarted GMRES(m) (Saad & Schultz 1986) for A x = b with x0 = 0; the
    operator is a callable apply_A(v, out) on device vectors; m <= 63.  arnoldi
    (library_arnoldi) may do the Arnoldi step for the same A in one library call.
    precond (a device vector s) right-preconditions with diag(s): the iteration runs
    on A diag(s) and x = diag(s) y, so the residual it minimises is A's own (Saad,
    Iterative Methods for Sparse Linear Systems, 2nd ed., 9.3.2).  Returns the number
    of iterations."""
    if not 1 <= restart <= 63:
        raise ValueError("gmres_device: restart must lie in [1, 63]")
    n = b.n
    V, w, r = _workspace(n, restart)
    x.fill_zero()
    bnorm = math.sqrt(_lib.dot(b, b))
    if bnorm == 0:
        return 0
    total = 0
    while total < maxiter:
        apply_A(x, w)
        _lib.lincomb(r, (1.0, -1.0), (b, w))
        beta = math.sqrt(_lib.dot(r, r))
        if beta <= rtol * bnorm:
            break
        _lib.lincomb(V[0], (1.0 / beta,), (r,))
        H = np.zeros((restart + 1, restart))
        g = np.zeros(restart + 1)
        g[0] = beta
        cs = np.zeros(restart)
        sn = np.zeros(restart)
        k_done = 0
        for j in range(restart):
            if arnoldi is not None:
                h, h2, ww = arnoldi(V, j, w) if precond is None else arnoldi(V, j, w, precond, r)
            else:
                if precond is None:
                    apply_A(V[j], w)
                else:
                    _lib.mul(r, precond, V[j])  # r is free until the next restart
                    apply_A(r, w)
                # classical Gram-Schmidt applied twice (CGS2: as stable as modified
                # Gram-Schmidt, Giraud et al. 2005), each pass one fused multi-dot and one
                # multi-axpy: two host synchronisations instead of j + 1
                basis = V[:j + 1]
                h = _lib.mdot(basis, w)
                _lib.maxpy(w, -h, basis)
                h2 = _lib.mdot(basis, w)
                _lib.maxpy(w, -h2, basis)
                ww = _lib.dot(w, w)
            H[:j + 1, j] = h + h2
            H[j + 1, j] = math.sqrt(ww)
            if H[j + 1, j] > 0:
                _lib.lincomb(V[j + 1], (1.0 / H[j + 1, j],), (w,))
            for i in range(j):  # apply previous Givens rotations
                a, c = H[i, j], H[i + 1, j]
                H[i, j], H[i + 1, j] = cs[i] * a + sn[i] * c, -sn[i] * a + cs[i] * c
            d = math.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = (1.0, 0.0) if d == 0 else (H[j, j] / d, H[j + 1, j] / d)
            H[j, j] = d
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            total += 1
            k_done = j + 1
            if abs(g[j + 1]) <= rtol * bnorm or total >= maxiter:
                break
        y = np.linalg.solve(np.triu(H[:k_done, :k_done]), g[:k_done]) if k_done else np.zeros(0)
        if k_done:
            if precond is None:
                _lib.maxpy(x, y, V[:k_done])
            else:
                r.fill_zero()
                _lib.maxpy(r, y, V[:k_done])
                _lib.mul(r, precond, r)
                _lib.lincomb(x, (1.0, 1.0), (x, r))
        if abs(g[k_done]) <= rtol * bnorm:
            break
    return total
