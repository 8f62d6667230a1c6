"""Device time integration (solve.jl:133-220): the plate diffusion workload
(diffusion_equation_on_a_square_plate.jl, 50 x 50 on [0, 2]^2, D = 1/9,
homogeneous Dirichlet, u0 = 50 for y <= 1) against the oracle's RHS stepped by
the same fixed-step scheme with the Dirichlet callback, against a tight reference
integration of the oracle's RHS, and against the exact double series
(diffusion_equation_on_a_square_plate.jl:79-94) at discretisation tolerance."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

import finitevolumemethod as FVM
from problems import oracle_for

pytestmark = pytest.mark.gpu


def plate_exact(x, y, t):
    """200/pi^2 sum_{m odd <= 49} sum_{n <= 50} (2/m) sin(m pi x/2) e^{-pi^2 m^2 t/36}
    (1 - cos(n pi/2))/n sin(n pi y/2) e^{-pi^2 n^2 t/36}."""
    s = np.zeros_like(x)
    for m in range(1, 50, 2):
        mt = 2 / m * np.sin(m * np.pi * x / 2) * np.exp(-np.pi ** 2 * m ** 2 * t / 36)
        for n in range(1, 51):
            nt = (1 - np.cos(n * np.pi / 2)) / n * np.sin(n * np.pi * y / 2) * np.exp(
                -np.pi ** 2 * n ** 2 * t / 36)
            s += mt * nt
    return 200 * s / np.pi ** 2


def plate(n=50, zero_boundary=False):
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, n, n, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    u0 = np.where(tri.y <= 1.0, 50.0, 0.0)
    if zero_boundary:
        u0[tri.boundary_nodes()] = 0.0
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1 / 9,
                          initial_condition=u0, final_time=0.5)
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=0, dpoly=[1 / 9, 0, 0, 0],
                                             fn=([0.0], [0.0], [0.0]))))
    return tri, prob, orc, u0


def test_fixed_step_trajectory_matches_oracle(gpu):
    """SSPRK3 with dt = 1e-3 and the Dirichlet callback after every step."""
    tri, prob, orc, u0 = plate()
    dt, nsteps = 1e-3, 500
    sol = FVM.solve(prob, FVM.SSPRK3(dt=dt))
    assert len(sol.u) == nsteps + 1
    u, t = u0.copy(), 0.0
    for k in range(1, nsteps + 1):
        k1 = orc.rhs(u, t)
        u1 = u + dt * k1
        u2 = 0.75 * u + 0.25 * u1 + 0.25 * dt * orc.rhs(u1, t + dt)
        u = 1 / 3 * u + 2 / 3 * u2 + 2 / 3 * dt * orc.rhs(u2, t + 0.5 * dt)
        t = t + dt
        u = orc.update_dirichlet(u, t)
        if k in (1, 10, 100, 500):
            assert np.max(np.abs(sol.u[k] - u)) <= 1e-12 * 50, k
            assert sol.t[k] == pytest.approx(t, abs=1e-12)


def test_tsit5_matches_tight_reference_integration(gpu):
    tri, prob, orc, u0 = plate(zero_boundary=True)
    ts = np.arange(0, 0.5001, 0.05)
    sol = FVM.solve(prob, FVM.Tsit5(), saveat=0.05, abstol=1e-10, reltol=1e-10)
    ref = solve_ivp(lambda t, u: orc.rhs(u, t), (0.0, 0.5), u0, method="DOP853", t_eval=ts,
                    rtol=1e-12, atol=1e-12)
    assert ref.status == 0 and len(sol.u) == len(ts)
    np.testing.assert_allclose(sol.t, ts, atol=1e-12)
    for k in range(len(ts)):
        assert np.max(np.abs(sol.u[k] - ref.y[:, k])) <= 1e-6 * 50, k


def test_tsit5_matches_exact_series(gpu):
    """Default tolerances (abstol 1e-6, reltol 1e-3), saveat 0.05: relative L2 error
    below 5e-3 and pointwise below 0.6 at every saved time t > 0 (the semi-discrete
    error of this 50 x 50 mesh is 3.3e-3 / 0.38 at t = 0.05 and decays after)."""
    tri, prob, orc, u0 = plate()
    sol = FVM.solve(prob, FVM.Tsit5(), saveat=0.05)
    assert sol.retcode == "Success" and len(sol.u) == 11
    np.testing.assert_array_equal(sol.u[0], u0)
    for t, u in zip(sol.t[1:], sol.u[1:]):
        e = plate_exact(tri.x, tri.y, t)
        assert np.linalg.norm(u - e) <= 5e-3 * np.linalg.norm(e), t
        assert np.max(np.abs(u - e)) <= 0.6, t


def test_dirichlet_nodes_stay_fixed_after_steps(gpu):
    tri, prob, orc, u0 = plate(n=20)
    sol = FVM.solve(prob, FVM.Tsit5(), saveat=0.1)
    bn = tri.boundary_nodes()
    for u in sol.u[1:]:
        assert np.all(u[bn] == 0.0)


def test_steady_problem_solves_to_a_zero_residual(gpu):
    """SteadyFVMProblem(FVMProblem): Newton-GMRES on fvm_jvp to F(u) = 0 for a
    nonlinear D(u) with a reaction term; the oracle's du vanishes at the answer."""
    import oracle as O  # noqa: F401
    from problems import oracle_for

    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 23, 21, jitter=0.2, seed=12)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 0.1 + u ** 2,
                          source_function=lambda x, y, t, u, p: 1.0 - u,
                          initial_condition=np.zeros(tri.num_points()), final_time=np.inf)
    sol = FVM.solve(FVM.SteadyFVMProblem(prob))
    assert sol.retcode == "Success", sol.stats
    u = sol.u[-1]
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=0, dpoly=[0.1, 0, 1, 0], spoly=[1, -1, 0, 0],
                                             fn=([0.0], [0.0], [0.0]))))
    r = orc.rhs(u, 0.0)
    assert np.linalg.norm(r) <= 1e-7 * np.linalg.norm(orc.scale(u, 0.0)), sol.stats
    assert np.all(u[tri.boundary_nodes()] == 0.0) and u.max() > 0.05
