"""TR-BDF2 on the device (the reference tutorials' TRBDF2): Newton-Krylov on fvm_jvp
(fvm_system_jvp for systems; for the ODE templates, the template matrix).

Checks: the plate problem against a tight reference integration of the oracle's RHS
with the same Dirichlet semantics and against the exact series; a nonlinear problem
against the oracle integrated by scipy's Radau; the Brusselator system against its
exact solution (the tutorial's own comparison); the DiffusionEquation template
against its FVMProblem twin; and stiffness: on a fine mesh TR-BDF2 takes far fewer
steps than Tsit5 for the same tolerance."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

import finitevolumemethod as FVM
from problems import brusselator_system, nonlinear_problem
from test_gpu_solve import plate, plate_exact

pytestmark = pytest.mark.gpu


def test_trbdf2_matches_tight_reference_integration(gpu):
    tri, prob, orc, u0 = plate(zero_boundary=True)
    ts = np.arange(0, 0.5001, 0.05)
    sol = FVM.solve(prob, FVM.TRBDF2(), saveat=0.05, abstol=1e-8, reltol=1e-8)
    assert sol.retcode == "Success", sol.stats
    ref = solve_ivp(lambda t, u: orc.rhs(u, t), (0.0, 0.5), u0, method="DOP853", t_eval=ts,
                    rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(sol.t, ts, atol=1e-12)
    for k in range(len(ts)):
        assert np.max(np.abs(sol.u[k] - ref.y[:, k])) <= 2e-5 * 50, (k, sol.stats)


def test_trbdf2_matches_exact_series(gpu):
    tri, prob, orc, u0 = plate()
    sol = FVM.solve(prob, FVM.TRBDF2(), saveat=0.05)
    assert sol.retcode == "Success" and len(sol.u) == 11
    for t, u in zip(sol.t[1:], sol.u[1:]):
        e = plate_exact(tri.x, tri.y, t)
        assert np.linalg.norm(u - e) <= 1e-2 * np.linalg.norm(e), t
    bn = tri.boundary_nodes()
    assert all(np.all(u[bn] == 0.0) for u in sol.u[1:])


def test_trbdf2_nonlinear_matches_radau(gpu):
    prob, orc, u = nonlinear_problem(nx=14, ny=12)
    tri = prob.mesh.triangulation
    u0 = u.copy()
    u0[tri.boundary_nodes()] = 0.0  # the Dirichlet callback holds them at 0
    prob.initial_condition = u0
    prob.final_time = 0.3
    ts = [0.0, 0.1, 0.2, 0.3]
    sol = FVM.solve(prob, FVM.TRBDF2(), saveat=ts, abstol=1e-9, reltol=1e-9)
    assert sol.retcode == "Success", sol.stats
    ref = solve_ivp(lambda t, v: orc.rhs(v, t), (0.0, 0.3), u0, method="Radau", t_eval=ts,
                    rtol=1e-11, atol=1e-12)
    for k in range(len(ts)):
        assert np.max(np.abs(sol.u[k] - ref.y[:, k])) <= 1e-6, (k, sol.stats)


def test_trbdf2_system_tracks_brusselator_exact_solution(gpu):
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 30, 30, single_boundary=False)
    sys = brusselator_system(tri, final_time=1.0)
    sol = FVM.solve(sys, FVM.TRBDF2(), saveat=0.5, abstol=1e-7, reltol=1e-7)
    assert sol.retcode == "Success", sol.stats
    for t, U in zip(sol.t, sol.u):
        ex = np.stack([np.exp(-tri.x - tri.y - t / 2), np.exp(tri.x + tri.y + t / 2)])
        for v in range(2):
            assert np.linalg.norm(U[v] - ex[v]) <= 1e-3 * np.linalg.norm(ex[v]), (t, v)


def test_trbdf2_template_equals_problem_path(gpu):
    """DiffusionEquation (J v = A v on the template matrix) and the FVMProblem (fvm_jvp)
    give the same trajectory to the integration tolerance."""
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, 30, 30, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    u0 = np.where(tri.y <= 1.0, 50.0, 0.0)
    u0[tri.boundary_nodes()] = 0.0
    tmpl = FVM.DiffusionEquation(mesh, BCs, diffusion_function=lambda x, y, p: 1 / 9,
                                 initial_condition=u0, final_time=0.5)
    gen = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1 / 9,
                         initial_condition=u0, final_time=0.5)
    a = FVM.solve(gen, FVM.TRBDF2(), saveat=0.25, abstol=1e-9, reltol=1e-9)
    b = FVM.solve(tmpl, FVM.TRBDF2(), saveat=0.25, abstol=1e-9, reltol=1e-9)
    assert a.retcode == b.retcode == "Success"
    for ua, ub in zip(a.u, b.u):
        assert np.max(np.abs(ua - ub[:-1])) <= 1e-6 * 50


def test_trbdf2_takes_large_steps_on_stiff_problems(gpu):
    """150 x 150 plate (stiffness ~ 8 D / h^2 ~ 5000) to t = 0.5: TR-BDF2's steps are set
    by accuracy, Tsit5's by stability."""
    tri, prob, orc, u0 = plate(n=150, zero_boundary=True)
    imp = FVM.solve(prob, FVM.TRBDF2(), saveat=0.5)
    exp_ = FVM.solve(prob, FVM.Tsit5(), saveat=0.5)
    assert imp.retcode == "Success" and exp_.retcode == "Success"
    # measured: 75 TR-BDF2 steps against 712 Tsit5 steps (the start is resolved by both)
    assert imp.stats["naccept"] * 5 < exp_.stats["naccept"], (imp.stats, exp_.stats)
    assert np.max(np.abs(imp.u[-1] - exp_.u[-1])) <= 0.05


def test_affine_problems_run_in_one_launch_and_match_the_host_integrator(gpu):
    """f(u) = J u + c (constant D, Neumann flux or Dirichlet values constant): the
    Jacobian extracted from fvm_jvp probes equals the JVP, the one-workgroup TR-BDF2
    (fvm_trbdf2_affine) follows the host integrator step for step to the Krylov and
    Newton tolerances, and nonlinear or time-dependent problems stay on the host."""
    from finitevolumemethod import _lib, implicit

    tri, prob, orc, u0 = plate()
    assert implicit.is_affine(prob)
    A = implicit.affine_operator(prob)
    rp, col, val = A.csr
    import scipy.sparse as sp
    J = sp.csr_matrix((val, col, rp), shape=(A.n, A.n))
    rng = np.random.default_rng(3)
    ctx = prob.device_context()
    for _ in range(2):
        v = rng.standard_normal(A.n)
        jv = _lib.DeviceArray(A.n)
        ctx.jvp(_lib.DeviceArray.from_host(u0), _lib.DeviceArray.from_host(v), jv, 0.0)
        ref = jv.to_host()
        assert np.max(np.abs(J @ v - ref)) <= 1e-13 * np.max(np.abs(ref))
        f = FVM.fvm_eqs(np.zeros(A.n), v, prob, 0.0)
        assert np.max(np.abs(J @ v + A.d_c.to_host() - f)) <= 1e-13 * np.max(np.abs(f))
    a = FVM.solve(prob, FVM.TRBDF2(), saveat=0.05)
    b = FVM.solve(prob, FVM.TRBDF2(device_loop=False), saveat=0.05)
    assert a.stats.get("device_loop") and not b.stats.get("device_loop")
    assert a.retcode == b.retcode == "Success"
    assert abs(a.stats["naccept"] - b.stats["naccept"]) <= 2, (a.stats, b.stats)
    for x, y in zip(a.u, b.u):
        assert np.max(np.abs(x - y)) <= 1e-5 * 50
    # a Neumann problem (the 100 x 100 walkthrough's kind) at a smaller size
    L = 320.0
    tri = FVM.triangulate_rectangle(0.0, L, 0.0, L, 40, 40, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    prob = FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 2.0 + 0.0 * u, FVM.Neumann),
                          diffusion_function=lambda x, y, t, u, p: 2.0 + 0.0 * u,
                          initial_condition=np.where((0.4 * L <= tri.y) & (tri.y <= 0.6 * L), 1.0, 0.0),
                          final_time=500.0)
    a = FVM.solve(prob, FVM.TRBDF2(), saveat=100.0)
    b = FVM.solve(prob, FVM.TRBDF2(device_loop=False), saveat=100.0)
    assert a.stats.get("device_loop") and a.retcode == b.retcode == "Success"
    for x, y in zip(a.u, b.u):
        assert np.max(np.abs(x - y)) <= 1e-4 * max(1.0, np.max(np.abs(y)))
    # the Neumann value is the outward flux q.n: mass falls by 2 per unit length of the
    # boundary (4 L) per unit time
    V = mesh.cv_volumes
    for t, u in zip(a.t, a.u):
        expect = V @ a.u[0] - 2.0 * 4 * L * t
        assert abs(V @ u - expect) <= 1e-6 * max(abs(expect), V @ a.u[0])
    prob_nl, _, _ = nonlinear_problem(nx=10, ny=9)
    assert not implicit.is_affine(prob_nl)


def test_jacobian_diagonal_and_the_jacobi_preconditioned_krylov_solves(gpu):
    """diag(J) from one fvm_jvp per colour of a distance-2 colouring equals the
    oracle's complex-step derivative of each du_i in u_i (node by node, 1e-11 of the
    largest entry); TR-BDF2 with GMRES right-preconditioned by diag(W)^-1
    (fvm_arnoldi_step_scaled) reaches the unpreconditioned solution to the
    integrator's tolerance.  (The iteration savings grow with the mesh: at 3,240
    nodes the two take about the same number of Krylov iterations; bench.py's
    solve_trbdf2 row reports both at 10^6 nodes.)"""
    from finitevolumemethod import _lib
    from finitevolumemethod.solve import jacobian_diagonal

    prob, orc, u = nonlinear_problem(nx=22, ny=19)
    n = len(u)
    out = _lib.DeviceArray(n)
    jacobian_diagonal(prob)(_lib.DeviceArray.from_host(u), 0.0, out)
    d = out.to_host()
    h = 1e-30
    ref = np.empty(n)
    for i in range(n):
        z = u.astype(np.complex128)
        z[i] += 1j * h
        ref[i] = orc.rhs_cs(z).imag[i] / h
    assert np.max(np.abs(d - ref)) <= 1e-11 * np.max(np.abs(ref))
    assert np.count_nonzero(ref) > n // 2
    prob, _, _ = nonlinear_problem(nx=60, ny=54)
    prob.final_time = 0.05
    a = FVM.solve(prob, FVM.TRBDF2(), saveat=[0.05])
    b = FVM.solve(prob, FVM.TRBDF2(precond="none"), saveat=[0.05])
    assert a.retcode == b.retcode == "Success"
    assert a.stats["nkrylov"] <= 1.1 * b.stats["nkrylov"], (a.stats, b.stats)
    assert np.max(np.abs(a.u[-1] - b.u[-1])) <= 1e-3 * np.max(np.abs(b.u[-1]))
    with pytest.raises(ValueError):
        FVM.solve(prob, FVM.TRBDF2(precond="ilu"), saveat=[0.05])
