"""Differentiation rules for fvm_eqs in the form of ChainRulesCore's rrule / frule.

The reference differentiates fvm_eqs! with ForwardDiff dual numbers pushed through
its DiffCache (main_equations.jl:47-67, solve.jl:5); there a gradient costs one
dual evaluation per input direction.  Here the hand-written adjoint kernel
(fvm_vjp) gives the whole pullback of du = fvm_eqs(u, p, t) in one launch, and the
tangent kernel (fvm_jvp) the pushforward:

    du, pullback = rrule(fvm_eqs, u, p, t)
    _, u_bar, p_bar, _ = pullback(du_bar)      # u_bar = J_u^T du_bar
    du, du_dot = frule((NoTangent(), u_dot, NoTangent(), NoTangent()), fvm_eqs, u, p, t)

p is an FVMProblem (or its parameter object).  For a problem whose diffusion is a
NodalDiffusion field, p_bar is Tangent(log_d=J_logD^T du_bar), the cotangent of the
per-node log D field; otherwise NoTangent().  Arrays may be NumPy (results come back
as NumPy) or DeviceArray (results stay on the device).  The pullback evaluates the
adjoint at the u and log D field of the rrule call: both are copied when the rule is
built, so later changes to either do not reach it.
"""
from . import _lib
from .equations import _prob_of, fvm_eqs
from .forms import NodalDiffusion
from .system import FVMSystem


class NoTangent:
    """The cotangent of an argument that is not differentiated (ChainRulesCore.NoTangent)."""

    def __repr__(self):
        return "NoTangent()"

    def __eq__(self, other):
        return isinstance(other, NoTangent)

    def __hash__(self):
        return 0


class Tangent:
    """The structural cotangent of a parameter object (ChainRulesCore.Tangent): the
    differentiated fields as attributes."""

    def __init__(self, **fields):
        self.__dict__.update(fields)

    def __repr__(self):
        return f"Tangent({', '.join(sorted(self.__dict__))})"


def _device_copy(a):
    if isinstance(a, _lib.DeviceArray):
        c = _lib.DeviceArray(a.n)
        c.copy_from(a)
        return c
    return _lib.DeviceArray.from_host(a)


def _problem(f, p):
    if f is not fvm_eqs:
        raise TypeError("rrule / frule: only fvm_eqs has a rule here")
    prob = _prob_of(p)
    if isinstance(prob, FVMSystem) or not hasattr(prob, "device_context"):
        raise TypeError("rrule / frule: p must be an FVMProblem (systems have no adjoint kernels)")
    return prob


def rrule(f, u, p, t=0.0):
    """(du, pullback) of du = fvm_eqs(u, p, t)."""
    prob = _problem(f, p)
    ctx = prob.device_context()
    host = not isinstance(u, _lib.DeviceArray)
    d_u = _device_copy(u)
    d_du = _lib.DeviceArray(d_u.n)
    ctx.rhs(d_u, d_du, t)
    nodal = isinstance(prob.diffusion_function, NodalDiffusion)
    field = None
    if nodal:  # the log D field in use now, kept for the pullback
        cur = getattr(ctx, "_log_d_ref", None)
        field = _device_copy(cur) if cur is not None else _lib.DeviceArray.from_host(
            prob.diffusion_function.log_d)

    def pullback(du_bar):
        w = du_bar if isinstance(du_bar, _lib.DeviceArray) else _lib.DeviceArray.from_host(du_bar)
        if w.n != d_u.n:
            raise ValueError("pullback: du_bar has the wrong length")
        ub = _lib.DeviceArray(d_u.n)
        lb = _lib.DeviceArray(d_u.n) if nodal else None
        prev = getattr(ctx, "_log_d_ref", None)
        if nodal:
            ctx.set_log_d(field)
        try:
            ctx.vjp(d_u, w, ub, lb, t)
        finally:
            if nodal and prev is not None:
                ctx.set_log_d(prev)
        if host:
            ub = ub.to_host()
            lb = lb.to_host() if lb is not None else None
        p_bar = Tangent(log_d=lb) if nodal else NoTangent()
        return NoTangent(), ub, p_bar, NoTangent()

    return (d_du.to_host() if host else d_du), pullback


def frule(dargs, f, u, p, t=0.0):
    """(du, du_dot) of du = fvm_eqs(u, p, t) pushed forward along dargs =
    (f_dot, u_dot, p_dot, t_dot); only u_dot may be a tangent (the others NoTangent)."""
    prob = _problem(f, p)
    _, u_dot, p_dot, t_dot = dargs
    if not isinstance(p_dot, NoTangent) or not isinstance(t_dot, NoTangent):
        raise NotImplementedError("frule: tangents in p or t are not supported (u only)")
    ctx = prob.device_context()
    host = not isinstance(u, _lib.DeviceArray)
    d_u = u if not host else _lib.DeviceArray.from_host(u)
    d_v = u_dot if isinstance(u_dot, _lib.DeviceArray) else _lib.DeviceArray.from_host(u_dot)
    d_du, d_jv = _lib.DeviceArray(d_u.n), _lib.DeviceArray(d_u.n)
    ctx.rhs(d_u, d_du, t)
    ctx.jvp(d_u, d_v, d_jv, t)
    if host:
        return d_du.to_host(), d_jv.to_host()
    return d_du, d_jv
