"""Closures outside the traced forms, compiled at problem set-up (jit.py, hipRTC,
FVM_FLUX_JIT): fvm_eqs! against the oracle calling the same closures
(construct_flux_function's flux, problem.jl:425-440, and the source) at the 1e-12
bar of the RHS tests, on meshes with Dirichlet and Neumann boundaries and at two
times; the JVP (generated dD/du, dS/du) against central differences of the device
RHS and the VJP by the adjoint identity; a solve; the routing of closures."""
import numpy as np
import pytest

import finitevolumemethod as FVM
from finitevolumemethod import _lib
from finitevolumemethod.problem import physics_of
from problems import oracle_for
from test_gpu_rhs import assert_parity

pytestmark = pytest.mark.gpu

CASES = {
    # D(x, y, t, u), S(x, y, t, u)
    "exp_sin": (lambda x, y, t, u: np.exp(-u), lambda x, y, t, u: np.sin(u)),
    "rational": (lambda x, y, t, u: 1.0 / (1.0 + u ** 2), lambda x, y, t, u: u * (1 - u) / (1 + u * u)),
    "xy_t": (lambda x, y, t, u: (1.0 + x) * np.exp(-u) + 0.1 * np.sin(y * u) + 0.05 * t,
             lambda x, y, t, u: np.cos(x) * u - t * np.tanh(u)),
    "where_pow": (lambda x, y, t, u: np.where(u > 0.5, u ** 2, 0.25 + 0.0 * u) + 2.0 ** (u - 1.0),
                  lambda x, y, t, u: np.sqrt(1.0 + u * u) - 1.0),
}


def jit_problem(name, nx=31, ny=27, seed=5):
    D, S = CASES[name]
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, nx, ny, jitter=0.25, seed=seed, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(
        mesh, (lambda x, y, t, u, p: 0.3 + 0.0 * u, lambda x, y, t, u, p: 0.0 * u,
               lambda x, y, t, u, p: -0.2 + 0.0 * u, lambda x, y, t, u, p: 0.7 + 0.0 * u),
        (FVM.Neumann, FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet))
    u = np.random.default_rng(seed).uniform(0.0, 1.0, tri.num_points())
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: D(x, y, t, u),
                          source_function=lambda x, y, t, u, p: S(x, y, t, u),
                          initial_condition=u, final_time=1.0)

    def flux(x, y, t, a, b, g, T):
        Dv = D(x, y, t, a * x + b * y + g)
        return -Dv * a, -Dv * b

    orc = oracle_for(prob, dict(flux=flux, source=lambda x, y, t, u, n: S(x, y, t, u),
                                cond=lambda f, x, y, t, u: prob.conditions.functions[f](x, y, t, u)))
    return prob, orc, u


@pytest.mark.parametrize("name", sorted(CASES))
def test_rhs_matches_the_oracle(gpu, name):
    prob, orc, u = jit_problem(name)
    assert prob.device_context().flux_kind == _lib.FLUX_JIT
    for t in (0.0, 0.37):
        du = FVM.fvm_eqs(np.zeros_like(u), u, prob, t)
        assert_parity(du, orc.rhs(u, t), orc.scale(u, t))


@pytest.mark.parametrize("name", sorted(CASES))
def test_jvp_and_vjp(gpu, name):
    prob, orc, u = jit_problem(name)
    ctx = prob.device_context()
    n = len(u)
    rng = np.random.default_rng(7)
    t = 0.2
    v, w = rng.standard_normal(n), rng.standard_normal(n)
    jv, jtw = _lib.DeviceArray(n), _lib.DeviceArray(n)
    ctx.jvp(_lib.DeviceArray.from_host(u), _lib.DeviceArray.from_host(v), jv, t)
    ctx.vjp(_lib.DeviceArray.from_host(u), _lib.DeviceArray.from_host(w), jtw, None, t)
    Jv, JTw = jv.to_host(), jtw.to_host()
    h = 1e-6
    fd = (FVM.fvm_eqs(np.zeros(n), u + h * v, prob, t) - FVM.fvm_eqs(np.zeros(n), u - h * v, prob, t)) / (2 * h)
    scale = orc.scale(u, t) * (1 + np.abs(v))
    if name == "where_pow":  # the selection's jump at u_s = 0.5: compare away from it
        ok = np.abs(fd - Jv) <= 1e-6 * scale
        assert ok.mean() > 0.97
    else:
        assert np.all(np.abs(Jv - fd) <= 1e-6 * scale), np.max(np.abs(Jv - fd) / scale)
    assert abs(w @ Jv - JTw @ v) <= 1e-11 * np.sum(np.abs(w * Jv))


def test_solve_with_a_compiled_closure(gpu):
    """Tsit5 through the device loop on a compiled closure equals the host loop bit
    for bit, and the steady state of u_t = div(exp(-u) grad u) + sin(u)... reached
    by the explicit integrator stays bounded (the Dirichlet values 0 and 0.7)."""
    prob, orc, u = jit_problem("exp_sin", 20, 18)
    prob.final_time = 0.05
    a = FVM.solve(prob, FVM.Tsit5(), saveat=[0.0, 0.05])
    b = FVM.solve(prob, FVM.Tsit5(), saveat=[0.0, 0.05], callback=lambda *args: None)
    assert a.retcode == "Success" and b.retcode == "Success"
    assert np.array_equal(a.u[-1], b.u[-1])
    # against the oracle stepped by the same (fixed) Euler steps
    dt = 1e-4
    c = FVM.solve(prob, FVM.Euler(), dt=dt, saveat=[0.0, 0.05])
    ref = u.copy()
    for k in range(500):
        ref = orc.update_dirichlet(ref + dt * orc.rhs(ref, k * dt), (k + 1) * dt)
    assert np.max(np.abs(c.u[-1] - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_problems_outside_the_catalogue_are_compiled(gpu):
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 6, 6, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    u0 = np.zeros(tri.num_points())
    mk = lambda D, S=None: FVM.FVMProblem(mesh, BCs, diffusion_function=D, source_function=S,  # noqa: E731
                                          initial_condition=u0, final_time=1.0)
    ph = physics_of(mk(lambda x, y, t, u, p: np.exp(-u)))
    assert ph["flux_kind"] == _lib.FLUX_JIT and "fvm_jit_dD" in ph["jit_source"]
    ph = physics_of(mk(lambda x, y, t, u, p: 0.1 + u ** 2, lambda x, y, t, u, p: np.sin(u)))
    assert ph["flux_kind"] == _lib.FLUX_JIT
    ph = physics_of(mk(lambda x, y, t, u, p: 0.1 + u ** 2, lambda x, y, t, u, p: u * (1 - u)))
    assert ph["flux_kind"] == _lib.FLUX_D_POLY_U and "jit_source" not in ph
    with pytest.raises(FVM.UnsupportedClosure):
        physics_of(mk(lambda x, y, t, u, p: 1.0 if u > 0 else 2.0))
