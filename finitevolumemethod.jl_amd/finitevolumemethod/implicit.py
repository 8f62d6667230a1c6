"""TR-BDF2 for small problems whose right-hand side is affine in u, in one launch.

When the traced closures make fvm_eqs! affine in u and constant in t (a diffusion
coefficient that does not depend on u, a source s(x, y) + s1 u, conditions with
values constant in t and Dirichlet values independent of u), f(u) = J u + c with a
constant J.  The reference hands such problems to TRBDF2 with a KLU factorisation of
J on its jacobian_sparsity pattern (solve.jl:50-77, 167-183).  Here J is extracted
once per problem from fvm_jvp probes over a distance-2 colouring of that pattern (one
probe per colour, about a dozen on a triangle mesh; exact for an affine f), stored in
ELL form, and fvm_trbdf2_affine integrates the whole problem in one workgroup (no host
round trip per Krylov iteration).  Problems above 65,536 unknowns, with a user
callback, without save times or outside the affine forms keep the host-driven
integrator (solve.py).
"""
import ctypes

import numpy as np

from . import _lib
from .equations import csr_pattern

MAX_N = 65536


class AffineOp(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("width", ctypes.c_int32), ("d_col", _lib._vp), ("d_val", _lib._vp),
                ("d_diag", _lib._vp), ("d_ovf_rp", _lib._vp), ("d_ovf_col", _lib._vp), ("d_ovf_val", _lib._vp),
                ("d_c", _lib._vp), ("n_dirichlet", ctypes.c_int32), ("d_dir_nodes", _lib._vp),
                ("d_dir_values", _lib._vp)]


class TrbOpts(ctypes.Structure):
    _fields_ = [("t0", ctypes.c_double), ("tf", ctypes.c_double), ("dt", ctypes.c_double),
                ("abstol", ctypes.c_double), ("reltol", ctypes.c_double), ("maxiters", ctypes.c_int64),
                ("newton_kappa", ctypes.c_double), ("max_newton", ctypes.c_int32),
                ("krylov_restart", ctypes.c_int32), ("krylov_rtol", ctypes.c_double),
                ("smooth_est", ctypes.c_int32), ("n_save", ctypes.c_int64), ("save_t", _lib._dp),
                ("d_saves", _lib._vp)]


class TrbStats(ctypes.Structure):
    _fields_ = [("nf", ctypes.c_int64), ("njvp", ctypes.c_int64), ("naccept", ctypes.c_int64),
                ("nreject", ctypes.c_int64), ("nnewton", ctypes.c_int64), ("nkrylov", ctypes.c_int64),
                ("nnewton_fail", ctypes.c_int64), ("nsaved", ctypes.c_int64), ("retcode", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("t", ctypes.c_double)]


_RETCODES = {0: "Success", 1: "MaxIters", 2: "Unstable", 3: "ConvergenceFailure"}


def is_affine(prob):
    """Whether the problem's traced right-hand side is affine in u and constant in t."""
    from .problem import FVMProblem

    if not isinstance(prob, FVMProblem):
        return False
    ph = prob.device_context().physics
    fk = ph["flux_kind"]
    if fk == _lib.FLUX_D_POLY_U:
        if any(ph.get("dpoly", [0.0] * 4)[1:]):
            return False
    elif fk not in (_lib.FLUX_D_NODAL, _lib.FLUX_D_EDGE, _lib.FLUX_LINEAR):
        return False
    sp = ph.get("spoly", [0.0] * 4)
    if sp[2] != 0.0 or sp[3] != 0.0 or ph.get("st", 0.0) != 0.0:
        return False
    fn = ph["fn"]
    ct, lam = np.asarray(fn[2]), (np.asarray(fn[5]) if len(fn) > 5 else np.zeros(1))
    if np.any(ct != 0.0) or np.any(lam != 0.0):
        return False
    kind, fidx = ph["node_kind"], ph["node_fidx"]
    used = np.unique(fidx[kind == _lib.NODE_DIRICHLET])
    return not np.any(np.asarray(fn[1])[used] != 0.0)


def _colouring(rp, col, n):
    """Greedy distance-2 colouring of a symmetric pattern: columns that share a row
    get different colours (fvmh_colour_d2 in the host library)."""
    rp = np.ascontiguousarray(rp, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    colour = np.empty(n, np.int32)
    nc = _lib.host_lib().fvmh_colour_d2(n, _lib.ptr(rp, _lib._i64p), _lib.ptr(col, _lib._i32p),
                                        _lib.ptr(colour, _lib._i32p))
    if nc < 0:
        raise ValueError(_lib.host_lib().fvmh_last_error().decode())
    return colour.astype(np.int64)


class AffineOperator:
    """J (ELL form on the device), c = f(0) and the Dirichlet values of an affine
    problem, built once."""

    def __init__(self, prob):
        ctx = prob.device_context()
        tri = prob.mesh.triangulation
        n = tri.num_points()
        t0 = prob.initial_time
        rp, col = csr_pattern(tri)
        rp = rp.astype(np.int64)
        colour = _colouring(rp, col, n)
        rows = np.repeat(np.arange(n), np.diff(rp))
        val = np.zeros(len(col))
        zero = _lib.DeviceArray(n)
        zero.fill_zero()
        jv = _lib.DeviceArray(n)
        for c in range(int(colour.max()) + 1):
            ctx.jvp(zero, _lib.DeviceArray.from_host((colour == c).astype(np.float64)), jv, t0)
            Jv = jv.to_host()
            sel = colour[col] == c
            val[sel] = Jv[rows[sel]]
        self.n_colours = int(colour.max()) + 1
        f0 = _lib.DeviceArray(n)
        ctx.rhs(zero, f0, t0)
        # ELL: the off-diagonals of each row, padded with the row's own column
        diag = val[col == rows].copy() if np.all(np.bincount(rows[col == rows], minlength=n) == 1) else None
        if diag is None:
            raise ValueError("the pattern must hold every diagonal entry")
        off = col != rows
        cnt = np.bincount(rows[off], minlength=n)
        W = int(cnt.max()) if n else 0
        ecol = np.repeat(np.arange(n, dtype=np.int32)[None, :], W, axis=0)
        eval_ = np.zeros((W, n))
        slot = np.arange(len(col)) - rp[rows]  # position in the row, diagonal included
        # slot among the off-diagonals: subtract 1 after the diagonal
        dpos = np.zeros(n, np.int64)
        dpos[rows[~off]] = slot[~off]
        s_off = slot - (slot > dpos[rows])
        ecol[s_off[off], rows[off]] = col[off]
        eval_[s_off[off], rows[off]] = val[off]
        self.W = W
        self.d_col = _lib.DeviceArray.from_host(ecol.ravel(), np.int32)
        self.d_val = _lib.DeviceArray.from_host(eval_.ravel())
        self.d_diag = _lib.DeviceArray.from_host(diag)
        self.d_orp = _lib.DeviceArray.from_host(np.zeros(n + 1, np.int32), np.int32)
        self.d_c = f0
        kind = ctx.physics["node_kind"]
        dn = np.nonzero(kind == _lib.NODE_DIRICHLET)[0].astype(np.int32)
        probe = _lib.DeviceArray(n)
        probe.fill_zero()
        ctx.apply_dirichlet(probe, t0)  # the Dirichlet values (constant in t and u)
        self.dir_nodes = dn
        self.d_dir_nodes = _lib.DeviceArray.from_host(dn if len(dn) else np.zeros(1, np.int32), np.int32)
        self.d_dir_values = _lib.DeviceArray.from_host(probe.to_host()[dn] if len(dn) else np.zeros(1))
        self.n = n
        self.csr = (rp, col, val)

    def op(self):
        o = AffineOp()
        o.n, o.width = self.n, self.W
        o.d_col, o.d_val, o.d_diag = self.d_col.ptr, self.d_val.ptr, self.d_diag.ptr
        o.d_ovf_rp, o.d_ovf_col, o.d_ovf_val = self.d_orp.ptr, None, None
        o.d_c = self.d_c.ptr
        o.n_dirichlet = len(self.dir_nodes)
        o.d_dir_nodes, o.d_dir_values = self.d_dir_nodes.ptr, self.d_dir_values.ptr
        return o


def affine_operator(prob):
    op = getattr(prob, "_affine_op", None)
    if op is None:
        op = AffineOperator(prob)
        prob._affine_op = op
    return op


def eligible(prob, alg, saveat, user_callback):
    from .problem import FVMProblem

    return (isinstance(prob, FVMProblem) and getattr(alg, "device_loop", True) and user_callback is None
            and saveat is not None and len(prob.initial_condition) <= MAX_N and 1 <= alg.krylov_restart <= 30
            and is_affine(prob))


def trbdf2_affine(prob, alg, saveat, abstol, reltol, dt, maxiters, save_start):
    """solve(prob, TRBDF2(); saveat) through fvm_trbdf2_affine (see the module
    docstring); the same ODESolution as the host integrator."""
    from .solve import ODESolution, _initial_dt, _Saver

    A = affine_operator(prob)
    ctx = prob.device_context()
    n = A.n
    t0, tf = prob.initial_time, prob.final_time
    u = _lib.DeviceArray.from_host(np.asarray(prob.initial_condition, np.float64))
    fn = _lib.DeviceArray(n)
    nf = 1
    ctx.rhs(u, fn, t0)
    sol = ODESolution()
    saver = _Saver(saveat, t0, tf)
    while saver.k < len(saver.times) and saver.times[saver.k] <= t0 + 1e-14 * max(1, abs(t0)):
        sol.t.append(saver.times[saver.k])
        sol.u.append(u.to_host())
        saver.k += 1
    if dt is None:
        def f(x, out, tt):
            ctx.rhs(x, out, tt)
        dt = _initial_dt(f, u, fn, t0, tf, abstol, reltol, n, exponent=1 / 3)
        nf += 1
    times = np.ascontiguousarray(saver.times[saver.k:], np.float64)
    d_saves = _lib.DeviceArray(max(1, len(times) * n))
    o = TrbOpts()
    o.t0, o.tf, o.dt, o.abstol, o.reltol, o.maxiters = t0, tf, float(dt), abstol, reltol, int(maxiters)
    o.newton_kappa, o.max_newton = alg.newton_kappa, alg.max_newton
    o.krylov_restart, o.krylov_rtol, o.smooth_est = alg.krylov_restart, alg.krylov_rtol, int(alg.smooth_est)
    o.n_save, o.save_t, o.d_saves = len(times), _lib.ptr(times), d_saves.ptr
    st = TrbStats()
    opd = A.op()
    _lib.check(_lib.hip_lib().fvm_trbdf2_affine(ctypes.byref(opd), ctypes.byref(o), u.ptr, ctypes.byref(st), None),
               "fvm_trbdf2_affine")
    if st.nsaved:
        S = d_saves.to_host()[: st.nsaved * n].reshape(st.nsaved, n)
        for k in range(st.nsaved):
            sol.t.append(float(times[k]))
            sol.u.append(S[k].copy())
    sol.retcode = _RETCODES.get(st.retcode, "Failure")
    sol.stats = dict(nf=nf + st.nf, njvp=st.njvp, naccept=st.naccept, nreject=st.nreject, nnewton=st.nnewton,
                     nkrylov=st.nkrylov, nnewton_fail=st.nnewton_fail, device_loop=True,
                     jacobian_colours=A.n_colours)
    sol.final_device = u
    return sol
