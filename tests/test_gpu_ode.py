"""fvm_ode_solve: the explicit integrators run whole in the library (solve.jl:133-220
semantics: Dirichlet update after every accepted step, saveat) against the
host-language loop over the same kernels.  The two perform the same arithmetic in
the same order, so every saved state, time and counter must agree bit for bit;
the host loop itself is pinned against the oracle's RHS in test_gpu_solve.py."""
import ctypes

import numpy as np
import pytest

import finitevolumemethod as FVM
from finitevolumemethod import _lib
from finitevolumemethod.solve import fsal_ok, integrate
from problems import annulus_problem, brusselator_system, mixed_problem

pytestmark = pytest.mark.gpu


def plate(n=50):
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, n, n, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    u0 = np.where(tri.y <= 1.0, 50.0, 0.0)
    return FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1 / 9,
                          initial_condition=u0, final_time=0.5)


def host_loop(prob, alg, **kw):
    """The same integration through the host-language loop (no library operator)."""
    ctx = prob.device_context()
    cb = (lambda u, t: ctx.apply_dirichlet(u, t)) if prob.conditions.has_dirichlet_nodes() else None
    return integrate(ctx.rhs, prob.initial_condition, prob.initial_time, prob.final_time, alg,
                     step_callback=cb, fsal=fsal_ok(prob), **kw)


def assert_same(a, b, nf=True):
    assert a.retcode == b.retcode == "Success"
    if nf:
        assert a.stats == b.stats
    else:
        assert {k: v for k, v in a.stats.items() if k != "nf"} == {k: v for k, v in b.stats.items() if k != "nf"}
    assert a.t == b.t and len(a.u) == len(b.u)
    for k, (x, y) in enumerate(zip(a.u, b.u)):
        assert np.all(np.isfinite(x)), k
        assert np.array_equal(x, y), (k, a.t[k], np.max(np.abs(x - y)))


@pytest.mark.parametrize("saveat", [0.05, 0.001, [0.0, 0.013, 0.2, 0.5]])
def test_tsit5_plate_matches_host_loop(gpu, saveat):
    prob = plate()
    a = FVM.solve(prob, FVM.Tsit5(), saveat=saveat)
    b = host_loop(prob, FVM.Tsit5(), saveat=saveat)
    assert_same(a, b)
    assert a.stats["naccept"] > 50


@pytest.mark.parametrize("alg", [FVM.SSPRK3(dt=2e-4), FVM.Euler(dt=1e-4)], ids=["ssprk3", "euler"])
@pytest.mark.parametrize("variant", ["tutorial", "exp"])
def test_fixed_step_matches_host_loop(gpu, alg, variant):
    """The annulus with a time-dependent Dirichlet boundary (updated after every
    step) and with Neumann / Dudt boundaries, to t = 0.05 (250 or 500 steps)."""
    prob, _, _ = annulus_problem(variant, final_time=0.05)
    a = FVM.solve(prob, alg, saveat=0.01)
    b = host_loop(prob, alg, saveat=0.01)
    assert_same(a, b)
    assert len(a.u) == 6


def test_internal_conditions_match_host_loop(gpu):
    """mixed_problem (internal Dirichlet and Dudt nodes, t0 = 2) over a few steps:
    its Dirichlet functions move u by about -0.5 per update, so only a short run
    stays well posed."""
    prob, _, _ = mixed_problem()
    prob.final_time = 2.001
    a = FVM.solve(prob, FVM.SSPRK3(dt=1e-4), saveat=0.0005)
    b = host_loop(prob, FVM.SSPRK3(dt=1e-4), saveat=0.0005)
    assert_same(a, b)
    assert len(a.u) == 3


def test_tsit5_time_dependent_dirichlet_matches_host_loop(gpu):
    prob, _, _ = annulus_problem("tutorial", final_time=0.3)
    a = FVM.solve(prob, FVM.Tsit5(), saveat=0.1)
    b = host_loop(prob, FVM.Tsit5(), saveat=0.1)
    assert_same(a, b)
    tri = prob.mesh.triangulation
    inner = np.unique(tri.boundary_edges[tri.boundary_segments == 1])
    np.testing.assert_allclose(a.final_device.to_host()[inner], 50 * (1 - np.exp(-0.3 / 2)), rtol=1e-13)


def test_system_matches_host_loop(gpu):
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 20, 18, jitter=0.2, seed=2, single_boundary=False)
    sys = brusselator_system(tri, final_time=0.5)
    a = FVM.solve(sys, FVM.Tsit5(), saveat=0.1)
    ctx = sys.device_context()
    cb = (lambda u, t: ctx.apply_dirichlet(u, t)) if sys.has_dirichlet_nodes() else None
    b = integrate(ctx.rhs, sys.flat(sys.initial_condition), sys.initial_time, sys.final_time, FVM.Tsit5(),
                  saveat=0.1, step_callback=cb, fsal=fsal_ok(sys))
    b.u = [sys.unflat(v) for v in b.u]
    assert_same(a, b)
    assert a.u[0].shape == (2, tri.num_points())


def test_diffusion_template_matches_host_loop(gpu):
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, 30, 30, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    u0 = np.where(tri.y <= 1.0, 50.0, 0.0)
    prob = FVM.DiffusionEquation(mesh, BCs, diffusion_function=lambda x, y, p: 1 / 9,
                                 initial_condition=u0, final_time=0.5)
    a = FVM.solve(prob, FVM.Tsit5(), saveat=0.05)
    p = prob.problem
    b = integrate(lambda u, du, t: p.apply(u, du), prob._u0[:p.n], 0.0, 0.5, FVM.Tsit5(), saveat=0.05,
                  fsal=True)
    b.u = [np.concatenate([s, [1.0]]) for s in b.u]
    # a template this small runs its whole loop in one workgroup: the same arithmetic
    # per element, the error norm summed in the same order and the controller's power
    # the host's (fvm_ctrl.hpp): the same states bit for bit
    assert_same(a, b)


def test_large_diffusion_template_matches_host_loop_bit_for_bit(gpu):
    """Above the one-workgroup size (65,536 unknowns) the library loop is the
    multi-kernel one: every saved state, time and counter equals the host loop's."""
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, 280, 280, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    u0 = np.where(tri.y <= 1.0, 50.0, 0.0)
    prob = FVM.DiffusionEquation(mesh, BCs, diffusion_function=lambda x, y, p: 1 / 9,
                                 initial_condition=u0, final_time=2e-4)
    a = FVM.solve(prob, FVM.Tsit5(), saveat=1e-4)
    p = prob.problem
    b = integrate(lambda u, du, t: p.apply(u, du), prob._u0[:p.n], 0.0, 2e-4, FVM.Tsit5(), saveat=1e-4,
                  fsal=True)
    b.u = [np.concatenate([s, [1.0]]) for s in b.u]
    assert_same(a, b)


def test_paths_without_save_times_or_with_callbacks_stay_on_the_host_loop(gpu):
    """saveat=None saves every step and a user callback runs between steps: both
    keep the host loop and give the same states as the library at the common times."""
    prob = plate(20)
    every = FVM.solve(prob, FVM.Tsit5())
    assert len(every.u) == every.stats["naccept"] + 1
    seen = []
    cbk = FVM.solve(prob, FVM.Tsit5(), saveat=0.1, callback=lambda u, t: seen.append(t))
    lib = FVM.solve(prob, FVM.Tsit5(), saveat=0.1)
    assert len(seen) == lib.stats["naccept"]
    # a user callback may change u, so f(u) is evaluated again after each of its steps
    assert_same(cbk, lib, nf=False)
    assert cbk.stats["nf"] - lib.stats["nf"] == lib.stats["naccept"] - 1


def test_ode_solve_rejects_bad_options(gpu):
    prob = plate(10)
    ctx = prob.device_context()
    u = _lib.DeviceArray.from_host(prob.initial_condition)
    st = _lib.OdeStats()

    def call(**kw):
        o = _lib.OdeOpts()
        o.method, o.op_kind, o.op, o.n = _lib.ODE_TSIT5, _lib.OP_PROBLEM, ctx.handle, len(prob.initial_condition)
        o.t0, o.tf, o.abstol, o.reltol, o.maxiters = 0.0, 0.1, 1e-6, 1e-3, 1000
        for k, v in kw.items():
            setattr(o, k, v)
        return _lib.check(_lib.hip_lib().fvm_ode_solve(ctypes.byref(o), u.ptr, ctypes.byref(st), None),
                          "fvm_ode_solve")

    call()
    for bad in (dict(op=None), dict(op_kind=7), dict(method=9), dict(tf=-1.0), dict(abstol=0.0),
                dict(n_save=3), dict(method=_lib.ODE_EULER, dt=0.0)):
        with pytest.raises(FVM.FVMError):
            call(**bad)


def test_library_arnoldi_matches_host_gram_schmidt(gpu):
    """gmres_device with fvm_arnoldi_step (A = J and A = I - c J) gives the host
    Gram-Schmidt's iterates bit for bit."""
    from finitevolumemethod.linsolve import gmres_device, library_arnoldi

    prob = plate(40)
    ctx = prob.device_context()
    n = len(prob.initial_condition)
    rng = np.random.default_rng(5)
    u = _lib.DeviceArray.from_host(rng.random(n))
    b = _lib.DeviceArray.from_host(rng.standard_normal(n))
    for a, c in ((0.0, 1.0), (1.0, -0.05)):
        def apply(v, out, a=a, c=c):
            ctx.jvp(u, v, out, 0.1)
            if not (a == 0.0 and c == 1.0):
                _lib.lincomb(out, (a, c), (v, out))
        x1, x2 = _lib.DeviceArray(n), _lib.DeviceArray(n)
        k1 = gmres_device(apply, b, x1, rtol=1e-10, restart=20, maxiter=400)
        k2 = gmres_device(apply, b, x2, rtol=1e-10, restart=20, maxiter=400,
                          arnoldi=library_arnoldi(_lib.OP_PROBLEM, ctx.handle, u, 0.1, a, c))
        assert k1 == k2 and k1 > 20, (k1, k2)  # restarts included
        assert np.array_equal(x1.to_host(), x2.to_host())


def test_trbdf2_with_library_arnoldi_matches_host_loop(gpu):
    prob = plate(30)
    # device_loop=False: the host integrator with the library Arnoldi step (an affine
    # problem such as this one otherwise runs whole in fvm_trbdf2_affine)
    a = FVM.solve(prob, FVM.TRBDF2(device_loop=False, precond="none"), saveat=0.1)
    ctx = prob.device_context()
    cb = lambda u, t: ctx.apply_dirichlet(u, t)  # noqa: E731
    b = integrate(ctx.rhs, prob.initial_condition, prob.initial_time, prob.final_time, FVM.TRBDF2(),
                  saveat=0.1, step_callback=cb, jvp=lambda u, v, out, t: ctx.jvp(u, v, out, t))
    assert_same(a, b)
    assert a.stats["nkrylov"] > 50


def test_first_same_as_last_saves_one_evaluation_per_step(gpu):
    """Tsit5 takes f(u) after an accepted step from its last stage when the Dirichlet
    update cannot change u: constant Dirichlet values (from the second step on; the
    initial condition need not hold them) or none at all.  The states are the same
    bit for bit as with a fresh evaluation; time-dependent Dirichlet values keep it."""
    prob = plate(30)
    assert fsal_ok(prob)
    ctx = prob.device_context()
    cb = lambda u, t: ctx.apply_dirichlet(u, t)  # noqa: E731
    run = lambda fsal: integrate(ctx.rhs, prob.initial_condition, prob.initial_time, prob.final_time,  # noqa: E731
                                 FVM.Tsit5(), saveat=0.1, step_callback=cb, fsal=fsal)
    a, b = run(True), run(False)
    assert_same(a, b, nf=False)
    assert b.stats["nf"] - a.stats["nf"] == a.stats["naccept"] - 1
    # without Dirichlet nodes every accepted step reuses its last stage
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 25, 25, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    neu = FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Neumann),
                         diffusion_function=lambda x, y, t, u, p: 0.1 + 0.0 * u,
                         initial_condition=np.sin(3 * tri.x) * np.cos(2 * tri.y), final_time=0.2)
    assert fsal_ok(neu)
    c = FVM.solve(neu, FVM.Tsit5(), saveat=0.1)
    d = integrate(neu.device_context().rhs, neu.initial_condition, 0.0, 0.2, FVM.Tsit5(), saveat=0.1)
    assert_same(c, d, nf=False)
    assert d.stats["nf"] - c.stats["nf"] == c.stats["naccept"]
    # a Dirichlet boundary that moves with t does not qualify
    moving, _, _ = annulus_problem("tutorial", final_time=0.05)
    assert not fsal_ok(moving)
