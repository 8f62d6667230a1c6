"""The reference tutorials whose closures trace onto the device forms, against the
closed-form solutions the tutorials themselves compare with (their meshes come from
DelaunayTriangulation's refinement, absent here, so each test builds a mesh of the
same domain; the tolerances are discretisation-level, stated per test).

  diffusion_equation_in_a_wedge_with_mixed_boundary_conditions.jl: a Fourier-Bessel
      series (for f = 1 - r only its n = 0 terms are nonzero);
  helmholtz_equation_with_inhomogeneous_boundary_conditions.jl: steady, closed form;
  reaction_diffusion_equation_with_a_time_dependent_dirichlet_boundary_condition_on_a_disk.jl:
      exp(t) sqrt(I0(sqrt(2) r)), D = u, R = u (1 - u), du/dt = u on the boundary."""
import numpy as np
import pytest
import scipy.integrate as si
import scipy.special as sp

import finitevolumemethod as FVM
from problems import disk_triangulation, wedge_triangulation

pytestmark = pytest.mark.gpu


def test_wedge_mixed_conditions_series(gpu):
    """u_t = lap u on 0 < theta < pi/4, r < 1; du/dn = 0 on both straight edges,
    u = 0 on the arc, u0 = 1 - r: u = sum_m A_m exp(-z_m^2 t) J0(z_m r) with z_m the
    zeros of J0 and A_m = 2 / J1(z_m)^2 int_0^1 (1 - r) J0(z_m r) r dr (the
    tutorial's exact_solution, :107-156).  Relative L2 error below 1e-3 at t = 0.05
    and 0.1 on a 40-ring mesh (measured 1.8e-4, 2.6e-4)."""
    tri = wedge_triangulation(np.pi / 4, 40)
    mesh = FVM.FVMGeometry(tri)
    zero = lambda x, y, t, u, p: 0.0 * u  # noqa: E731
    BCs = FVM.BoundaryConditions(mesh, (zero, zero, zero), (FVM.Neumann, FVM.Dirichlet, FVM.Neumann))
    r = np.hypot(tri.x, tri.y)
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1.0 + 0.0 * u,
                          initial_condition=1.0 - r, final_time=0.1)
    sol = FVM.solve(prob, FVM.Tsit5(), saveat=[0.0, 0.05, 0.1], abstol=1e-9, reltol=1e-7)
    assert sol.retcode == "Success"
    z = sp.jn_zeros(0, 60)
    A = np.array([2 / sp.j1(zm) ** 2 * si.quad(lambda s: (1 - s) * sp.j0(zm * s) * s, 0, 1, limit=200)[0]
                  for zm in z])
    for t, u in zip(sol.t[1:], sol.u[1:]):
        e = np.sum(A[:, None] * np.exp(-z[:, None] ** 2 * t) * sp.j0(z[:, None] * r[None, :]), axis=0)
        print("wedge", t, np.linalg.norm(u - e) / np.linalg.norm(e))
        assert np.linalg.norm(u - e) <= 1e-3 * np.linalg.norm(e), (t, np.linalg.norm(u - e) / np.linalg.norm(e))


def test_helmholtz_inhomogeneous_neumann(gpu):
    """lap u + u = 0 on [-1, 1]^2, du/dn = 1 (q . n = -1), solved as a
    SteadyFVMProblem (D = 1, S = u): u = -(cos(x+1) + cos(1-x) + cos(y+1) + cos(1-y))
    / sin 2 (the tutorial's exact_solution, :80-82).  80 x 80 grid, relative L2
    error below 2e-4 (measured 5.8e-5).  The route taken is the affine problem's linear
    solve with the BiCGStab fallback."""
    tri = FVM.triangulate_rectangle(-1.0, 1.0, -1.0, 1.0, 80, 80, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: -1.0 + 0.0 * u, FVM.Neumann)
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1.0 + 0.0 * u,
                          source_function=lambda x, y, t, u, p: u,
                          initial_condition=np.zeros(tri.num_points()), final_time=np.inf)
    sol = FVM.solve(FVM.SteadyFVMProblem(prob))
    assert sol.retcode == "Success", sol.stats
    # affine in u, but the reaction term makes -diag(V) J indefinite: CG stops short and the
    # linear solve falls back to BiCGStab on the extracted Jacobian, with no Newton step
    assert sol.stats["linear_solve"]["method"].startswith("BiCGStab") and sol.stats["newton_iters"] == 0, sol.stats
    x, y = tri.x, tri.y
    e = -(np.cos(x + 1) + np.cos(1 - x) + np.cos(y + 1) + np.cos(1 - y)) / np.sin(2)
    u = sol.u[-1]
    print("helmholtz", np.linalg.norm(u - e) / np.linalg.norm(e), sol.stats)
    assert np.linalg.norm(u - e) <= 2e-4 * np.linalg.norm(e), np.linalg.norm(u - e) / np.linalg.norm(e)


def test_disk_time_dependent_boundary(gpu):
    """u_t = div(u grad u) + u (1 - u) on the unit disk, du/dt = u on the boundary
    circle, u0 = sqrt(I0(sqrt(2) r)): u = exp(t) sqrt(I0(sqrt(2) r)) (the tutorial's
    exact_solution, :82-85).  30-ring mesh to t = 0.1; relative error below 2e-4
    (measured 6.2e-5)."""
    tri = disk_triangulation(nr=30)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: u, FVM.Dudt)
    r = np.hypot(tri.x, tri.y)
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: u,
                          source_function=lambda x, y, t, u, p: u * (1 - u),
                          initial_condition=np.sqrt(sp.i0(np.sqrt(2) * r)), final_time=0.1)
    sol = FVM.solve(prob, FVM.Tsit5(), saveat=[0.0, 0.05, 0.1], abstol=1e-10, reltol=1e-8)
    assert sol.retcode == "Success"
    bn = tri.boundary_nodes()
    for t, u in zip(sol.t[1:], sol.u[1:]):
        e = np.exp(t) * np.sqrt(sp.i0(np.sqrt(2) * r))
        print("disk", t, np.max(np.abs(u - e) / e))
        assert np.max(np.abs(u - e) / e) <= 2e-4, (t, np.max(np.abs(u - e) / e))
        # the boundary follows du/dt = u exactly up to the integrator's tolerance
        np.testing.assert_allclose(u[bn], np.exp(t) * sol.u[0][bn], rtol=1e-6)


def test_porous_fisher_travelling_wave(gpu):
    """porous_fisher_equation_and_travelling_waves.jl:40-75: u_t = div(D u grad u) +
    lam u (1 - u) on [0, 3] x [0, 40], u = 1 at y = 0, u = 0 at y = 40, zero flux on the
    sides, u0 = 0; D = 0.9, lam = 0.99.  The sharp-front travelling wave
    u = 1 - exp(a z), z = y - s t - z0 <= 0, has s = sqrt(lam D / 2) and a = sqrt(lam /
    (2 D)) (substituting the profile into the travelling-wave ODE; the tutorial's
    variable names c and c_min hold these two the other way round).  The solution does
    not depend on x, so a coarse x spacing suffices (6 x 161 grid).  The front (u = 1/2)
    moves at s within 1 % between t = 15 and t = 25 (measured 0.6 %), and behind it the
    profile matches the exact one within 5e-3 (measured 1.3e-3)."""
    Dd, lam = 0.9, 0.99
    tri = FVM.triangulate_rectangle(0.0, 3.0, 0.0, 40.0, 6, 161, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    one = lambda x, y, t, u, p: 1.0 + 0.0 * u  # noqa: E731
    zero = lambda x, y, t, u, p: 0.0 * u  # noqa: E731
    BCs = FVM.BoundaryConditions(mesh, (one, zero, zero, zero),
                                 (FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet, FVM.Neumann))
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, D: D * u, diffusion_parameters=Dd,
                          source_function=lambda x, y, t, u, l: l * u * (1 - u), source_parameters=lam,
                          initial_condition=np.zeros(tri.num_points()), final_time=25.0)
    sol = FVM.solve(prob, FVM.Tsit5(), saveat=[15.0, 25.0])
    assert sol.retcode == "Success"
    col = np.nonzero(np.isclose(tri.x, 1.8))[0]
    col = col[np.argsort(tri.y[col])]
    y = tri.y[col]

    def front(u):
        v = u[col]
        k = np.nonzero(v < 0.5)[0][0]
        return y[k - 1] + (v[k - 1] - 0.5) / (v[k - 1] - v[k]) * (y[k] - y[k - 1])

    s, a = np.sqrt(lam * Dd / 2), np.sqrt(lam / (2 * Dd))
    speed = (front(sol.u[1]) - front(sol.u[0])) / 10.0
    print("porous fisher speed", speed, "exact", s)
    assert abs(speed - s) <= 0.01 * s
    y0 = front(sol.u[1]) + np.log(2) / a  # where the exact profile reaches 0
    behind = (y < y0 - 0.5) & (y > y0 - 8)
    err = np.max(np.abs(sol.u[1][col][behind] - (1 - np.exp(a * (y[behind] - y0)))))
    print("porous fisher profile", err)
    assert err <= 5e-3


def test_laplace_internal_dirichlet_steady_equals_template(gpu):
    """laplaces_equation_with_internal_dirichlet_conditions.jl:47-112: lap u = 0 on
    [0, 1]^2 with u = 0 (bottom), 100 y (sides), 100 (top) and internal Dirichlet nodes
    u = 0 on x = 1/2, y <= 2/5.  The SteadyFVMProblem (Newton-Krylov on fvm_eqs!) and the
    LaplacesEquation template (device CG) give the same solution to the solvers'
    tolerances (1e-5 absolute on values up to 100; measured 2.8e-7); the internal nodes
    hold 0 and the maximum principle holds."""
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 41, 41, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    bcs = (lambda x, y, t, u, p: 0.0 * y, lambda x, y, t, u, p: 100.0 * y,
           lambda x, y, t, u, p: 100.0 + 0.0 * y, lambda x, y, t, u, p: 100.0 * y)
    BCs = FVM.BoundaryConditions(mesh, bcs, (FVM.Dirichlet,) * 4)
    line = np.nonzero((tri.x == 0.5) & (tri.y <= 0.4))[0]
    ICs = FVM.InternalConditions((lambda x, y, t, u, p: 0.0 * y,), dirichlet_nodes={int(v): 0 for v in line})
    u0 = np.where((tri.x == 0.5) & (tri.y <= 0.4), 0.0, 100 * tri.y)
    prob = FVM.FVMProblem(mesh, BCs, ICs, diffusion_function=lambda x, y, t, u, p: 1.0 + 0.0 * u,
                          initial_condition=u0, final_time=np.inf)
    s1 = FVM.solve(FVM.SteadyFVMProblem(prob))
    assert s1.retcode == "Success", s1.stats
    tmpl = FVM.LaplacesEquation(mesh, BCs, ICs, diffusion_function=lambda x, y, p: 1.0)
    s2 = FVM.solve(tmpl, rtol=1e-12)
    u1, u2 = s1.u[-1], s2.u
    print("laplace steady vs template", np.max(np.abs(u1 - u2)))
    assert np.max(np.abs(u1 - u2)) <= 1e-5
    assert np.all(u1[line] == 0.0)
    assert u1.min() >= -1e-9 and u1.max() <= 100 + 1e-9


def test_poisson_template_equals_steady_problem(gpu):
    """poissons_equation.jl:92-163: laplace u = -sin(pi x) sin(pi y) on a 100 x 100 grid
    of [0, 1]^2 with u = 0 on the boundary, solved by the PoissonsEquation template and as
    a SteadyFVMProblem (the source's sign flipped, :119-130): both are within rtol 1e-4 of
    sin(pi x) sin(pi y) / (2 pi^2) and of each other (the tutorial's two @test lines)."""
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 100, 100, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * x, FVM.Dirichlet)
    src = lambda x, y, p: -np.sin(np.pi * x) * np.sin(np.pi * y)  # noqa: E731
    sol = FVM.solve(FVM.PoissonsEquation(mesh, BCs, source_function=src))
    fvm_sol = FVM.solve(FVM.SteadyFVMProblem(FVM.FVMProblem(
        mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1.0 + 0.0 * u,
        source_function=lambda x, y, t, u, p: -src(x, y, p) + 0.0 * u,
        initial_condition=np.zeros(tri.num_points()), final_time=np.inf)))
    assert sol.retcode == "Success" and fvm_sol.retcode == "Success", fvm_sol.stats
    exact = np.sin(np.pi * tri.x) * np.sin(np.pi * tri.y) / (2 * np.pi ** 2)
    u1, u2 = sol.u, fvm_sol.u[-1]
    print("poisson template vs steady", np.linalg.norm(u1 - u2) / np.linalg.norm(u1), fvm_sol.stats)
    assert np.linalg.norm(u1 - exact) <= 1e-4 * np.linalg.norm(exact)
    assert np.linalg.norm(u2 - u1) <= 1e-4 * np.linalg.norm(u1)


def test_laplace_walkthrough_template_and_steady_problem(gpu):
    """laplaces_equation.jl:148-222: div[(x + 1)(y + 2) grad u] = 0 on [0, 5]^2 (100 x 100
    grid, four boundary segments), u = 0 at x = 0, u = 5 at x = 5, zero flux on y = 0, 5;
    exact u = 5 log(1 + x) / log 6.  The LaplacesEquation template (device CG) and the
    SteadyFVMProblem (Newton-GMRES on fvm_jvp, Jacobi-preconditioned and plain) match
    it to the tutorial's rtol 1e-3 (:194) and each other to the solvers' tolerances;
    the preconditioner cuts the Krylov iterations (D varies 42-fold over the domain).
    By default the affine steady problem is one linear solve on its extracted Jacobian
    (no Newton step) to the steady tolerances (|F| <= 1e-8 + 1e-8 |F0|), which agrees
    with the template's solution to 1e-6 like the Newton runs."""
    tri = FVM.triangulate_rectangle(0.0, 5.0, 0.0, 5.0, 100, 100, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    zero = lambda x, y, t, u, p: 0.0 * x  # noqa: E731
    five = lambda x, y, t, u, p: 5.0 + 0.0 * x  # noqa: E731
    BCs = FVM.BoundaryConditions(mesh, (zero, five, zero, zero),
                                 (FVM.Neumann, FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet))
    exact = 5.0 * np.log1p(tri.x) / np.log(6.0)
    lap = FVM.LaplacesEquation(mesh, BCs, diffusion_function=lambda x, y, p: (x + 1) * (y + 2))
    s_t = FVM.solve(lap)
    assert np.linalg.norm(s_t.u - exact) <= 1e-3 * np.linalg.norm(exact)
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: (x + 1) * (y + 2) + 0.0 * u,
                          initial_condition=np.zeros(tri.num_points()), final_time=np.inf)
    sols = {pc: FVM.solve(FVM.SteadyFVMProblem(prob), precond=pc, linear="newton") for pc in ("jacobi", "none")}
    for pc, s in sols.items():
        assert s.retcode == "Success" and s.stats["precond"] == pc, s.stats
        u = s.u[-1]
        assert np.linalg.norm(u - exact) <= 1e-3 * np.linalg.norm(exact), pc
        assert np.linalg.norm(u - s_t.u) <= 1e-6 * np.linalg.norm(s_t.u), pc
    assert sols["jacobi"].stats["nkrylov"] < sols["none"].stats["nkrylov"], {k: v.stats for k, v in sols.items()}
    lin = FVM.solve(FVM.SteadyFVMProblem(prob))
    assert lin.retcode == "Success" and lin.stats["newton_iters"] == 0 and "linear_solve" in lin.stats, lin.stats
    assert np.linalg.norm(lin.u[-1] - exact) <= 1e-3 * np.linalg.norm(exact)
    assert np.linalg.norm(lin.u[-1] - s_t.u) <= 1e-6 * np.linalg.norm(s_t.u)
    with pytest.raises(ValueError):
        FVM.solve(FVM.SteadyFVMProblem(prob), precond="ilu")
    with pytest.raises(ValueError):
        FVM.solve(FVM.SteadyFVMProblem(prob), linear="lu")


def test_capacitor_walkthrough_template_and_steady_problem(gpu):
    """poissons_equation.jl:185-382, the parallel-plate capacitor on [0, 10]^2 (here a
    51 x 51 grid on which the plates are grid lines): internal Dirichlet nodes V = -1 on
    2 <= x <= 8, y = 3 and V = 1 on y = 7, eps = 1 + (eps0 - 1)(1 + erf(r / 4)) / 2 and the
    source -Q exp(-r^2 / 2) / (2 pi eps0) of the distance r to the plates, Dirichlet 0 at
    the bottom and top, zero flux at the sides.  The PoissonsEquation template and the
    SteadyFVMProblem (one linear solve on the extracted Jacobian by default, Newton-GMRES
    on request) agree to 1e-6 and hold the plate and boundary values exactly."""
    from scipy.special import erf

    def dist(x, y):
        dx = np.asarray(x) - np.clip(x, 2.0, 8.0)
        return np.minimum(np.hypot(dx, np.asarray(y) - 3.0), np.hypot(dx, np.asarray(y) - 7.0))

    eps0, Q = 8.8541878128e-12, 1.0e-6
    diel = lambda x, y, p: 1 + (eps0 - 1) * (1 + erf(dist(x, y) / 4.0)) / 2  # noqa: E731
    src = lambda x, y, p: -(Q / (2 * np.pi) * np.exp(-dist(x, y) ** 2 / 2)) / eps0  # noqa: E731
    tri = FVM.triangulate_rectangle(0.0, 10.0, 0.0, 10.0, 51, 51, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    zero = lambda x, y, t, u, p: 0.0 * x  # noqa: E731
    BCs = FVM.BoundaryConditions(mesh, (zero, zero, zero, zero), (FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet, FVM.Neumann))
    on = (tri.x >= 2.0 - 1e-9) & (tri.x <= 8.0 + 1e-9)
    lower = np.nonzero(on & (np.abs(tri.y - 3.0) < 1e-9))[0]
    upper = np.nonzero(on & (np.abs(tri.y - 7.0) < 1e-9))[0]
    assert len(lower) == len(upper) == 31
    ICs = FVM.InternalConditions((lambda x, y, t, u, p: -1.0 + 0.0 * x, lambda x, y, t, u, p: 1.0 + 0.0 * x),
                                 dirichlet_nodes={**{int(i): 0 for i in lower}, **{int(i): 1 for i in upper}})
    tmpl = FVM.solve(FVM.PoissonsEquation(mesh, BCs, ICs, diffusion_function=diel, source_function=src))
    prob = FVM.FVMProblem(mesh, BCs, ICs, diffusion_function=lambda x, y, t, u, p: diel(x, y, None) + 0.0 * u,
                          source_function=lambda x, y, t, u, p: -src(x, y, None) + 0.0 * u,
                          initial_condition=np.zeros(tri.num_points()), final_time=np.inf)
    for linear in ("auto", "newton"):
        s = FVM.solve(FVM.SteadyFVMProblem(prob), linear=linear)
        assert s.retcode == "Success", (linear, s.stats)
        assert ("linear_solve" in s.stats) == (linear == "auto"), s.stats
        u = s.u[-1]
        assert np.linalg.norm(u - tmpl.u) <= 1e-6 * np.linalg.norm(tmpl.u), linear
        assert np.all(u[lower] == -1.0) and np.all(u[upper] == 1.0)
        bottom_top = np.nonzero((tri.y == 0.0) | (tri.y == 10.0))[0]
        assert np.all(u[bottom_top] == 0.0)
    assert np.all(tmpl.u[lower] == -1.0) and np.all(tmpl.u[upper] == 1.0)


def test_mean_exit_time_walkthrough_template_and_steady_problem(gpu):
    """mean_exit_time.jl:89-215 on a ring mesh of its domain (workloads.met_problems):
    the MeanExitTimeProblem template and the SteadyFVMProblem of D grad^2 T = -1 agree
    (the reference asserts rtol 1e-2 at :206, here 1e-6), T = 0 at the absorbing
    boundaries and the two point holes, T > 0 elsewhere, at the scale of the reference's
    figure."""
    from finitevolumemethod import workloads as WL

    tri, met, steady = WL.met_problems()
    a = FVM.solve(met)
    s = FVM.solve(steady)
    assert s.retcode == "Success" and "linear_solve" in s.stats, s.stats
    T = s.u[-1]
    assert np.linalg.norm(T - a.u) <= 1e-6 * np.linalg.norm(a.u)
    assert T[0] == 0.0 and T[1] == 0.0
    seg = tri.boundary_segments
    absorbing = np.unique(tri.boundary_edges[seg >= 1])
    assert np.all(T[absorbing] == 0.0)
    free = np.setdiff1d(np.arange(tri.num_points()), np.concatenate([absorbing, [0, 1]]))
    assert np.all(T[free] > 0.0)
    assert 1e3 < T.max() < 1e5  # the scale of the reference's figure (contours 0:1000:15000, :147)
