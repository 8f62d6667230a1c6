"""GPU Jacobian-vector and vector-Jacobian products of fvm_eqs!.

JVP: checked against complex-step differentiation of the CPU oracle (machine-
precision derivatives, Squire & Trapp 1998) to 1e-11 relative.  VJP: checked
against the JVP by the adjoint identity w . (J v) = (J^T w) . v for random v, and
against central finite differences of the GPU RHS with step 1e-6 (bar 1e-6)."""
import numpy as np
import pytest

import finitevolumemethod as FVM
from problems import dxy_problem, heat_convection_problem, mixed_problem, nodal_problem, nonlinear_problem

pytestmark = pytest.mark.gpu


def _dev(a):
    return FVM.DeviceArray.from_host(np.asarray(a, np.float64))


def jvp(prob, u, v, t=0.0):
    out = FVM.DeviceArray(len(u))
    prob.device_context().jvp(_dev(u), _dev(v), out, t)
    return out.to_host()


def vjp(prob, u, w, t=0.0, with_logd=False):
    ub = FVM.DeviceArray(len(u))
    lb = FVM.DeviceArray(len(u)) if with_logd else None
    prob.device_context().vjp(_dev(u), _dev(w), ub, lb, t)
    return ub.to_host(), (lb.to_host() if with_logd else None)


@pytest.mark.parametrize("builder", [nonlinear_problem, nodal_problem])
def test_jvp_matches_complex_step(gpu, builder):
    out = builder()
    prob, orc, u = out[:3]
    logd = out[3] if len(out) > 3 else None
    v = np.random.default_rng(5).standard_normal(len(u))
    h = 1e-30
    du_c = orc.rhs_cs(u + 1j * h * v, logd, 0.0)
    ref = du_c.imag / h
    got = jvp(prob, u, v)
    assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_logd_derivative_matches_complex_step(gpu):
    prob, orc, u, logd = nodal_problem()
    z = np.random.default_rng(6).standard_normal(len(u))
    w = np.random.default_rng(7).standard_normal(len(u))
    h = 1e-30
    ref = orc.rhs_cs(u.astype(complex), logd + 1j * h * z, 0.0).imag / h  # J_logD z
    _, lb = vjp(prob, u, w, with_logd=True)
    lhs = np.dot(lb, z)
    rhs = np.dot(w, ref)
    assert abs(lhs - rhs) <= 1e-11 * np.sum(np.abs(w * ref))


@pytest.mark.parametrize("builder", [nonlinear_problem, nodal_problem, mixed_problem, dxy_problem])
def test_vjp_adjoint_identity(gpu, builder):
    out = builder()
    prob, orc, u = out[:3]
    rng = np.random.default_rng(8)
    t = prob.initial_time
    w = rng.standard_normal(len(u))
    ub, _ = vjp(prob, u, w, t)
    for _ in range(3):
        v = rng.standard_normal(len(u))
        Jv = jvp(prob, u, v, t)
        lhs, rhs = np.dot(w, Jv), np.dot(ub, v)
        assert abs(lhs - rhs) <= 1e-11 * (np.sum(np.abs(w * Jv)) + 1e-300)


def test_vjp_neumann_boundaries(gpu):
    prob, orc, u0 = heat_convection_problem(40)
    u = u0 + np.random.default_rng(1).random(len(u0))
    rng = np.random.default_rng(2)
    w = rng.standard_normal(len(u))
    ub, _ = vjp(prob, u, w)
    v = rng.standard_normal(len(u))
    Jv = jvp(prob, u, v)
    assert abs(np.dot(w, Jv) - np.dot(ub, v)) <= 1e-11 * np.sum(np.abs(w * Jv))


def test_vjp_against_central_differences(gpu):
    """Central differences, step 1e-6, agree with the VJP to 1e-6."""
    prob, orc, u, logd = nodal_problem(40, 36)
    rng = np.random.default_rng(3)
    w = rng.standard_normal(len(u))
    ub, lb = vjp(prob, u, w, with_logd=True)
    ctx = prob.device_context()
    du_p, du_m = FVM.DeviceArray(len(u)), FVM.DeviceArray(len(u))
    h = 1e-6
    for _ in range(4):
        v = rng.standard_normal(len(u))
        ctx.rhs(_dev(u + h * v), du_p, 0.0)
        ctx.rhs(_dev(u - h * v), du_m, 0.0)
        fd = np.dot(w, (du_p.to_host() - du_m.to_host()) / (2 * h))
        assert abs(fd - np.dot(ub, v)) <= 1e-6 * max(1.0, abs(fd))
        z = rng.standard_normal(len(u))
        d_lp, d_lm = _dev(logd + h * z), _dev(logd - h * z)
        ctx.set_log_d(d_lp)
        ctx.rhs(_dev(u), du_p, 0.0)
        ctx.set_log_d(d_lm)
        ctx.rhs(_dev(u), du_m, 0.0)
        ctx.set_log_d(_dev(logd))
        fd = np.dot(w, (du_p.to_host() - du_m.to_host()) / (2 * h))
        assert abs(fd - np.dot(lb, z)) <= 1e-6 * max(1.0, abs(fd))
