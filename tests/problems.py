"""Problem builders shared by the tests: each returns the product problem and the
oracle problem with the same mesh, conditions and closures.

The closures mirror the reference's own test problems (test/test_functions.jl:
example_diffusion_problem, example_heat_convection_problem, example_problem) and
the workloads the benchmark uses (nonlinear D(u) = 0.1 + u^2 with S = u(1 - u),
a nodal log-diffusivity field, D(x, y) = exp(0.5 sin 4 pi x cos 4 pi y))."""
import numpy as np

import finitevolumemethod as FVM
from oracle import OracleProblem, OracleSystem, flatten_conditions, merge_conditions


def oracle_for(prob, physics, tri=None):
    tri = tri or prob.mesh.triangulation
    kind, fidx, bneu = prob.conditions.flatten(tri)
    return OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                         kind, fidx, bneu, physics)


def nonlinear_problem(nx=30, ny=26, seed=3, jitter=0.25):
    """Throughput workload: jittered unit square, homogeneous Dirichlet,
    D(u) = 0.1 + u^2, S = u (1 - u), u ~ U[0, 1)."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, nx, ny, jitter=jitter, seed=seed)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    u = np.random.default_rng(seed + 1).random(tri.num_points())
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 0.1 + u ** 2,
                          source_function=lambda x, y, t, u, p: u * (1 - u),
                          initial_condition=u, final_time=1.0)
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=0, dpoly=[0.1, 0, 1, 0],
                                             spoly=[0, 1, -1, 0], fn=([0.0], [0.0], [0.0]))))
    return prob, orc, u


def diffusion_problem(n=25):
    """example_diffusion_problem (test_functions.jl:315-328): [0,2]^2, D = 1/9,
    homogeneous Dirichlet, u0 = 50 for y <= 1."""
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, n, n, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    u0 = np.where(tri.y <= 1.0, 50.0, 0.0)
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1 / 9,
                          initial_condition=u0, final_time=0.5)
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=0, dpoly=[1 / 9, 0, 0, 0],
                                             fn=([0.0], [0.0], [0.0]))))
    return prob, orc, u0


HEAT = dict(L=1.0, k=237.0, T0=10.0, Tinf=10.0, alpha=80.0e-6, q=10.0, h=25.0)


def heat_convection_problem(n=200):
    """example_heat_convection_problem (test_functions.jl:346-374): four Neumann
    walls, flux q = -alpha grad u given as a flux function."""
    P = HEAT
    tri = FVM.triangulate_rectangle(0, P["L"], 0, P["L"], n, n, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    bot = lambda x, y, t, T, p: -p["alpha"] * p["q"] / p["k"]  # noqa: E731
    right = lambda x, y, t, T, p: 0.0 * T  # noqa: E731
    top = lambda x, y, t, T, p: -p["alpha"] * p["h"] / p["k"] * (p["Tinf"] - T)  # noqa: E731
    left = lambda x, y, t, T, p: 0.0 * T  # noqa: E731
    BCs = FVM.BoundaryConditions(mesh, (bot, right, top, left), (FVM.Neumann,) * 4,
                                 parameters=(P, None, P, None))
    flux = lambda x, y, t, a, b, g, p: (-p["alpha"] * a, -p["alpha"] * b)  # noqa: E731
    u0 = np.full(tri.num_points(), P["T0"])
    prob = FVM.FVMProblem(mesh, BCs, flux_function=flux, flux_parameters=dict(alpha=P["alpha"]),
                          initial_condition=u0, final_time=2000.0)
    c_top = -P["alpha"] * P["h"] / P["k"]
    fn = ([-P["alpha"] * P["q"] / P["k"], 0.0, c_top * P["Tinf"], 0.0],
          [0.0, 0.0, -c_top, 0.0], [0.0] * 4)
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=0, dpoly=[P["alpha"], 0, 0, 0], fn=fn)))
    return prob, orc, u0


def mixed_problem(nx=12, ny=19, seed=5):
    """example_problem-like (test_functions.jl:86-135) on [0,2]x[0,5] with four
    segments (Dirichlet, Neumann, Dudt, Neumann), internal Dirichlet and Dudt nodes,
    D(u) = 1 + 0.5 u, S = u + 1.5, all condition functions affine."""
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 5.0, nx, ny, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    f1 = lambda x, y, t, u, p: x * y + u - p  # noqa: E731  Dirichlet (bottom)
    f2 = lambda x, y, t, u, p: u + p - t  # noqa: E731       Neumann (right)
    f3 = lambda x, y, t, u, p: 2 * u + 0.25 * t - p  # noqa: E731  Dudt (top)
    f4 = lambda x, y, t, u, p: y - x  # noqa: E731           Neumann (left)
    BCs = FVM.BoundaryConditions(mesh, (f1, f2, f3, f4),
                                 (FVM.Dirichlet, FVM.Neumann, FVM.Dudt, FVM.Neumann),
                                 parameters=(0.5, 0.2, 0.3, 0.4))
    idir = {6 + k * nx: 0 for k in range(2, ny - 2)}
    idudt = {3 + 3 * nx: 1, 4 + 5 * nx: 1}
    ICs = FVM.InternalConditions((lambda x, y, t, u, p: x + y + t + u + p,
                                  lambda x, y, t, u, p: -u + 3 * t),
                                 dirichlet_nodes=idir, dudt_nodes=idudt, parameters=(0.29, None))
    u0 = np.random.default_rng(seed).random(tri.num_points())
    prob = FVM.FVMProblem(mesh, BCs, ICs, diffusion_function=lambda x, y, t, u, p: 1 + 0.5 * u,
                          source_function=lambda x, y, t, u, p: u + p, source_parameters=1.5,
                          initial_condition=u0, initial_time=2.0, final_time=5.0)
    # merged function list: (internal 0, internal 1, f1, f2, f3, f4)
    c0 = [0.29, 0.0, -0.5, 0.2, -0.3, 0.0]
    cu = [1.0, -1.0, 1.0, 1.0, 2.0, 0.0]
    ct = [1.0, 3.0, 0.0, -1.0, 0.25, 0.0]
    cx = [1.0, 0.0, 0.0, 0.0, 0.0, -1.0]
    cy = [1.0, 0.0, 0.0, 0.0, 0.0, 1.0]
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=0, dpoly=[1.0, 0.5, 0, 0],
                                             spoly=[1.5, 1.0, 0, 0], fn=(c0, cu, ct, cx, cy))))
    # f1 = x*y + u - p has an x*y term: give the oracle the exact closure instead
    orc_cb = oracle_for(prob, dict(
        flux=lambda x, y, t, a, b, g, T: (-(1 + 0.5 * (a * x + b * y + g)) * a,
                                          -(1 + 0.5 * (a * x + b * y + g)) * b),
        source=lambda x, y, t, u, n: u + 1.5,
        cond=lambda f, x, y, t, u: prob.conditions.functions[f](x, y, t, u)))
    return prob, orc_cb, u0


def nodal_problem(nx=28, ny=24, seed=11):
    """Per-node log-diffusivity field (the differentiation workload), D on the
    boundary and a homogeneous Dirichlet boundary."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, nx, ny, jitter=0.25, seed=seed)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    rng = np.random.default_rng(seed)
    logd = np.log(0.1) + 0.3 * rng.standard_normal(tri.num_points())
    u = rng.random(tri.num_points())
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=FVM.NodalDiffusion(logd),
                          initial_condition=u, final_time=0.1)
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=1, logD=logd, fn=([0.0], [0.0], [0.0]))))
    return prob, orc, u, logd


def dxy_problem(n=30, seed=13):
    """D(x, y) = exp(0.5 sin 4 pi x cos 4 pi y) on a jittered square with a Neumann
    bottom / right and Dirichlet top / left boundary, S = 0.3 - u^2."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, n, n, jitter=0.2, seed=seed,
                                 single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    D = lambda x, y, t, u, p: np.exp(0.5 * np.sin(4 * np.pi * x) * np.cos(4 * np.pi * y))  # noqa: E731
    BCs = FVM.BoundaryConditions(
        mesh, (lambda x, y, t, u, p: 0.5 + 0.0 * u, lambda x, y, t, u, p: -0.25 * u,
               lambda x, y, t, u, p: 0.0 * u, lambda x, y, t, u, p: 1.0 + 0.0 * u),
        (FVM.Neumann, FVM.Neumann, FVM.Dirichlet, FVM.Dirichlet))
    u = np.random.default_rng(seed).random(tri.num_points())
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=D,
                          source_function=lambda x, y, t, u, p: 0.3 - u ** 2,
                          initial_condition=u, final_time=1.0)
    fn = ([0.5, 0.0, 0.0, 1.0], [0.0, -0.25, 0.0, 0.0], [0.0] * 4)
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=2, dscale=1.0, spoly=[0.3, 0, -1, 0], fn=fn)))
    return prob, orc, u


def disk_triangulation(nr=12, R=1.0):
    """Rings of 6k points at radius R k / nr around the origin, Delaunay-triangulated."""
    pts = [(0.0, 0.0)]
    for k in range(1, nr + 1):
        r = R * k / nr
        m = 6 * k
        th = 2 * np.pi * np.arange(m) / m + 0.1 * k
        pts += list(zip(r * np.cos(th), r * np.sin(th)))
    P = np.array(pts)
    return FVM.triangulate(P[:, 0], P[:, 1])


def template_oracle(prob):
    """The oracle over a template problem's mesh and flattened conditions (its
    template_assemble needs no flux closure)."""
    tri = prob.mesh.triangulation
    kind, fidx, bneu = prob.conditions.flatten(tri)
    return OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                         kind, fidx, bneu, dict(flux=lambda *a: (0.0, 0.0)))


def system_conditions(tri, per_var_types):
    """Flattened per-variable conditions of an FVMSystem (merge_problem_conditions,
    problem.jl:378-410): variable v's functions follow those of variables < v."""
    N = tri.num_points()
    nv = len(per_var_types)
    kind = np.zeros((nv, N), np.int8)
    fidx = np.full((nv, N), -1, np.int32)
    bn = np.full((nv, len(tri.boundary_edges)), -1, np.int32)
    off = 0
    for v, types in enumerate(per_var_types):
        ne, ce, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments, types)
        k, f, b = flatten_conditions(N, tri.boundary_edges, ne, dn, dd)
        kind[v] = k
        fidx[v] = np.where(k != 0, f + off, -1)
        bn[v] = np.where(b >= 0, b + off, -1)
        off += len(types)
    return kind, fidx, bn


def brusselator_oracle(tri):
    """The Brusselator tutorial system (reaction_diffusion_brusselator_system_of_pdes.jl):
    exact solution Phi = exp(-x - y - t/2), Psi = exp(x + y + t/2)."""
    kind, fidx, bn = system_conditions(tri, [["Neumann", "Neumann", "Dirichlet", "Neumann"],
                                           ["Dirichlet", "Neumann", "Neumann", "Dirichlet"]])
    fs = [lambda x, y, t: -0.25 * np.exp(-x - t / 2), lambda x, y, t: 0.25 * np.exp(-1 - y - t / 2),
          lambda x, y, t: np.exp(-1 - x - t / 2), lambda x, y, t: -0.25 * np.exp(-y - t / 2),
          lambda x, y, t: np.exp(x + t / 2), lambda x, y, t: -0.25 * np.exp(1 + y + t / 2),
          lambda x, y, t: -0.25 * np.exp(1 + x + t / 2), lambda x, y, t: np.exp(y + t / 2)]
    return OracleSystem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind, fidx, bn,
                        lambda x, y, t, a, b, g: (np.array([-a[0] / 4, -a[1] / 4]), np.array([-b[0] / 4, -b[1] / 4])),
                        lambda v, x, y, t, u: u[0] ** 2 * u[1] - 2 * u[0] if v == 0 else -u[0] ** 2 * u[1] + u[0],
                        lambda f, x, y, t, u: fs[f](x, y, t))


BRUSSELATOR_PHI = (lambda x, y, t, u, p: -1 / 4 * np.exp(-x - t / 2),
                   lambda x, y, t, u, p: 1 / 4 * np.exp(-1 - y - t / 2),
                   lambda x, y, t, u, p: np.exp(-1 - x - t / 2),
                   lambda x, y, t, u, p: -1 / 4 * np.exp(-y - t / 2))
BRUSSELATOR_PSI = (lambda x, y, t, u, p: np.exp(x + t / 2),
                   lambda x, y, t, u, p: -1 / 4 * np.exp(1 + y + t / 2),
                   lambda x, y, t, u, p: -1 / 4 * np.exp(1 + x + t / 2),
                   lambda x, y, t, u, p: np.exp(y + t / 2))


def brusselator_system(tri, final_time=5.0, tile_triangles=0):
    """The Brusselator tutorial's FVMSystem, built through the public API."""
    mesh = FVM.FVMGeometry(tri, tile_triangles=tile_triangles)
    phi_BCs = FVM.BoundaryConditions(mesh, BRUSSELATOR_PHI,
                                     (FVM.Neumann, FVM.Neumann, FVM.Dirichlet, FVM.Neumann))
    psi_BCs = FVM.BoundaryConditions(mesh, BRUSSELATOR_PSI,
                                     (FVM.Dirichlet, FVM.Neumann, FVM.Neumann, FVM.Dirichlet))
    phi = FVM.FVMProblem(mesh, phi_BCs, flux_function=lambda x, y, t, a, b, g, p: (-a[0] / 4, -b[0] / 4),
                         source_function=lambda x, y, t, u, p: u[0] ** 2 * u[1] - 2 * u[0],
                         initial_condition=np.exp(-tri.x - tri.y), final_time=final_time)
    psi = FVM.FVMProblem(mesh, psi_BCs, flux_function=lambda x, y, t, a, b, g, p: (-a[1] / 4, -b[1] / 4),
                         source_function=lambda x, y, t, u, p: -u[0] ** 2 * u[1] + u[0],
                         initial_condition=np.exp(tri.x + tri.y), final_time=final_time)
    return FVM.FVMSystem(phi, psi)


def gray_scott_system(n=200, final_time=6000.0, tile_triangles=0):
    """The Gray-Scott tutorial's FVMSystem ([-1, 1]^2, n x n grid, zero-flux boundary)."""
    tri = FVM.triangulate_rectangle(-1.0, 1.0, -1.0, 1.0, n, n, single_boundary=True)
    mesh = FVM.FVMGeometry(tri, tile_triangles=tile_triangles)
    bc = lambda x, y, t, u, p: 0.0 * u[0] * 0.0 * u[1]  # noqa: E731
    r2 = tri.x ** 2 + tri.y ** 2
    u_prob = FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, bc, FVM.Neumann),
                            flux_function=lambda x, y, t, a, b, g, e: (-a[0] * e, -b[0] * e),
                            flux_parameters=0.00002,
                            source_function=lambda x, y, t, u, b: b * (1 - u[0]) - u[0] * u[1] ** 2,
                            source_parameters=0.04, initial_condition=1 - np.exp(-80 * r2),
                            final_time=final_time)
    v_prob = FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, bc, FVM.Neumann),
                            flux_function=lambda x, y, t, a, b, g, e: (-a[1] * e, -b[1] * e),
                            flux_parameters=0.00001,
                            source_function=lambda x, y, t, u, d: -d * u[1] + u[0] * u[1] ** 2,
                            source_parameters=0.1, initial_condition=np.exp(-80 * r2),
                            final_time=final_time)
    return FVM.FVMSystem(u_prob, v_prob)


def gray_scott_oracle(tri):
    kind, fidx, bn = system_conditions(tri, [["Neumann"], ["Neumann"]])
    return OracleSystem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind, fidx, bn,
                        lambda x, y, t, a, b, g: (np.array([-a[0] * 0.00002, -a[1] * 0.00001]),
                                                  np.array([-b[0] * 0.00002, -b[1] * 0.00001])),
                        lambda v, x, y, t, u: (0.04 * (1 - u[0]) - u[0] * u[1] ** 2 if v == 0
                                               else -0.1 * u[1] + u[0] * u[1] ** 2),
                        lambda f, x, y, t, u: 0.0)


def annulus_triangulation(R1=0.2, R2=1.0, nr=14, seed=0):
    """An annulus R1 <= r <= R2: concentric rings of points, Delaunay, the triangles of
    the hole removed; boundary segment 0 is the outer circle, 1 the inner (the
    domain on the left of every boundary edge)."""
    pts = []
    for k in range(nr + 1):
        r = R1 + (R2 - R1) * k / nr
        m = max(12, int(round(2 * np.pi * r / ((R2 - R1) / nr))))
        th = 2 * np.pi * np.arange(m) / m + 0.37 * k
        pts += list(zip(r * np.cos(th), r * np.sin(th)))
    P = np.array(pts)
    full = FVM.triangulate(P[:, 0], P[:, 1])
    T = full.triangles
    cx = P[T, 0].mean(1)
    cy = P[T, 1].mean(1)
    T = T[np.hypot(cx, cy) > R1]
    count = {}
    for a, b, c in T:
        for e in ((a, b), (b, c), (c, a)):
            count[e] = count.get(e, 0) + 1
    bedges = [e for e in count if (e[1], e[0]) not in count]
    bedges = np.array(sorted(bedges), np.int32)
    mid = np.hypot(P[bedges, 0].mean(1), P[bedges, 1].mean(1))
    seg = (mid < 0.5 * (R1 + R2)).astype(np.int32)
    return FVM.Triangulation(P[:, 0], P[:, 1], T, bedges, seg)


def annulus_problem(variant="tutorial", nr=14, final_time=2.0):
    """diffusion_equation_on_an_annulus.jl:54-78 on the annulus 0.2 <= r <= 1: D = 1,
    outer boundary Neumann zero, inner boundary Dirichlet 50 (1 - exp(-t/2)), the
    tutorial's three-Gaussian initial condition.  variant "exp" exercises the
    other condition kinds with an exp(lam t) term: outer Neumann
    0.25 exp(-x - t/2), inner Dudt y exp(t/2) - u/4."""
    tri = annulus_triangulation(nr=nr)
    mesh = FVM.FVMGeometry(tri)
    if variant == "tutorial":
        fns = (lambda x, y, t, u, p: 0.0 * u, lambda x, y, t, u, p: 50 * (1 - np.exp(-t / 2)) + 0.0 * u)
        types = (FVM.Neumann, FVM.Dirichlet)
    else:
        fns = (lambda x, y, t, u, p: 0.25 * np.exp(-x - t / 2) + 0.0 * u,
               lambda x, y, t, u, p: y * np.exp(0.5 * t) - 0.25 * u)
        types = (FVM.Neumann, FVM.Dudt)
    BCs = FVM.BoundaryConditions(mesh, fns, types)
    x, y = tri.x, tri.y
    u0 = (10 * np.exp(-25 * ((x + 0.5) ** 2 + (y + 0.5) ** 2))
          - 5 * np.exp(-50 * ((x + 0.3) ** 2 + (y + 0.5) ** 2))
          - 10 * np.exp(-45 * ((x - 0.5) ** 2 + (y - 0.5) ** 2)))
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1.0 + 0.0 * u,
                          initial_condition=u0, final_time=final_time)
    orc = oracle_for(prob, dict(
        flux=lambda x, y, t, a, b, g, T: (-a, -b),
        cond=lambda f, x, y, t, u: prob.conditions.functions[f](x, y, t, u)))
    return prob, orc, u0


def wedge_triangulation(alpha=np.pi / 4, nr=40):
    """The sector 0 <= theta <= alpha, r <= 1: rings of points (about as many per
    ring as its arc length allows at spacing 1/nr), Delaunay; boundary segments
    0: theta = 0, 1: the arc r = 1, 2: theta = alpha (the points on both straight
    edges are exactly collinear)."""
    c, s = np.cos(alpha), np.sin(alpha)
    if abs(alpha - np.pi / 4) < 1e-15:
        s = c  # y = x exactly on the upper edge
    pts = [(0.0, 0.0)]
    for k in range(1, nr + 1):
        r = k / nr
        m = max(1, int(np.ceil(alpha * r * nr)))
        for j in range(m + 1):
            th = alpha * j / m
            if j == 0:
                pts.append((r, 0.0))
            elif j == m:
                pts.append((r * c, r * s))
            else:
                pts.append((r * np.cos(th), r * np.sin(th)))
    P = np.array(pts)
    full = FVM.triangulate(P[:, 0], P[:, 1])
    be = full.boundary_edges
    mx, my = P[be, 0].mean(1), P[be, 1].mean(1)
    th = np.arctan2(my, mx)
    seg = np.where(np.abs(my) < 1e-12, 0, np.where(np.abs(th - alpha) < 1e-9, 2, 1)).astype(np.int32)
    return FVM.Triangulation(P[:, 0], P[:, 1], full.triangles, be, seg)
