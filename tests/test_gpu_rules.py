"""rrule / frule of fvm_eqs (the ChainRulesCore-style API over fvm_vjp / fvm_jvp).

The pullback's u cotangent and the pushforward are checked against the oracle's
complex-step derivative (J v) and against each other by the adjoint identity; the
log-D cotangent against the oracle's complex step in log D and against central
differences of the device RHS (step 1e-6, bar 1e-6)."""
import numpy as np
import pytest

import finitevolumemethod as FVM
from finitevolumemethod.rules import NoTangent, Tangent, frule, rrule
from problems import nodal_problem, nonlinear_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("builder", [nonlinear_problem, nodal_problem])
def test_rrule_and_frule_against_the_oracle(gpu, builder):
    out = builder()
    prob, orc, u = out[:3]
    logd = out[3] if len(out) > 3 else None
    rng = np.random.default_rng(31)
    du, pb = rrule(FVM.fvm_eqs, u, prob, 0.0)
    ref = orc.rhs(u, 0.0) if logd is None else orc.rhs_cs(u.astype(complex), logd, 0.0).real
    assert np.max(np.abs(du - ref)) <= 1e-12 * np.max(np.abs(ref))
    w = rng.standard_normal(len(u))
    f_bar, u_bar, p_bar, t_bar = pb(w)
    assert f_bar == NoTangent() and t_bar == NoTangent()
    for _ in range(2):
        v = rng.standard_normal(len(u))
        du2, jv = frule((NoTangent(), v, NoTangent(), NoTangent()), FVM.fvm_eqs, u, prob, 0.0)
        assert np.array_equal(du2, du)
        cs = orc.rhs_cs(u + 1j * 1e-30 * v, logd, 0.0).imag / 1e-30
        assert np.max(np.abs(jv - cs)) <= 1e-11 * np.max(np.abs(cs))
        assert abs(np.dot(w, jv) - np.dot(u_bar, v)) <= 1e-11 * np.sum(np.abs(w * jv))
    if logd is None:
        assert p_bar == NoTangent()
    else:
        assert isinstance(p_bar, Tangent)
        z = rng.standard_normal(len(u))
        cs = orc.rhs_cs(u.astype(complex), logd + 1j * 1e-30 * z, 0.0).imag / 1e-30
        assert abs(np.dot(p_bar.log_d, z) - np.dot(w, cs)) <= 1e-11 * np.sum(np.abs(w * cs))


def test_rrule_on_device_arrays_and_a_later_field_change(gpu):
    prob, orc, u, logd = nodal_problem()
    ctx = prob.device_context()
    rng = np.random.default_rng(32)
    w = rng.standard_normal(len(u))
    du_h, pb_h = rrule(FVM.fvm_eqs, u, prob)
    d_u = FVM.DeviceArray.from_host(u)
    du_d, pb_d = rrule(FVM.fvm_eqs, d_u, prob)
    assert isinstance(du_d, FVM.DeviceArray) and np.array_equal(du_d.to_host(), du_h)
    _, ub_d, pbar_d, _ = pb_d(FVM.DeviceArray.from_host(w))
    _, ub_h, pbar_h, _ = pb_h(w)
    assert np.array_equal(ub_d.to_host(), ub_h) and np.array_equal(pbar_d.log_d.to_host(), pbar_h.log_d)
    # change the field and u after building the rule: the pullback keeps the old ones
    ctx.set_log_d(FVM.DeviceArray.from_host(logd + 0.5))
    d_u.upload(u + 1.0)
    _, ub2, pbar2, _ = pb_d(FVM.DeviceArray.from_host(w))
    assert np.array_equal(ub2.to_host(), ub_h) and np.array_equal(pbar2.log_d.to_host(), pbar_h.log_d)
    # central differences in log D through the device RHS
    h = 1e-6
    z = rng.standard_normal(len(u))
    dp, dm = FVM.DeviceArray(len(u)), FVM.DeviceArray(len(u))
    ctx.set_log_d(FVM.DeviceArray.from_host(logd + h * z))
    ctx.rhs(FVM.DeviceArray.from_host(u), dp)
    ctx.set_log_d(FVM.DeviceArray.from_host(logd - h * z))
    ctx.rhs(FVM.DeviceArray.from_host(u), dm)
    ctx.set_log_d(FVM.DeviceArray.from_host(logd))
    fd = np.dot(w, (dp.to_host() - dm.to_host()) / (2 * h))
    assert abs(fd - np.dot(pbar_h.log_d, z)) <= 1e-6 * max(1.0, abs(fd))


def test_rules_reject_other_functions_and_systems(gpu):
    prob, orc, u = nonlinear_problem(8, 8)
    with pytest.raises(TypeError):
        rrule(np.sin, u, prob)
    with pytest.raises(NotImplementedError):
        frule((NoTangent(), u, Tangent(x=1), NoTangent()), FVM.fvm_eqs, u, prob)
