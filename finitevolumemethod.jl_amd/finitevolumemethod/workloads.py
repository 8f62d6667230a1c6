"""The benchmark workloads, built one way for the benchmark and the tests.

Each builder returns problems of this package with seeded synthetic data:

  throughput_problem      FVMProblem, 2000 x 2000 jittered Delaunay mesh on [0, 1]^2
                          (4,000,000 nodes, 7,992,002 triangles), D(u) = 0.1 + u^2,
                          S(u) = u (1 - u), homogeneous Dirichlet, u ~ U[0, 1)
  template_problems       PoissonsEquation and MeanExitTimeProblem on a 1000 x 1000
                          jittered mesh, D(x, y) = exp(0.5 sin 4 pi x cos 4 pi y),
                          f ~ U[-1, 1] (Poisson), homogeneous Dirichlet
  differentiation_problem FVMProblem with a per-node log-diffusivity field on a
                          1000 x 1000 jittered mesh: log D ~ N(log 0.1, 0.3^2),
                          u ~ U[0, 1), cotangent w ~ N(0, 1), homogeneous Dirichlet
  ensemble_members        256 diffusion problems on [0, 2]^2 with nx, ny drawn
                          independently as round(exp(U[log 20, log 300])), each
                          jittered with its own seed, D_p ~ LogNormal(log 0.1, 0.5),
                          u0 = 50 where y <= 1 (mid-height), else 0, homogeneous
                          Dirichlet

Interior points are jittered by U(-h/4, h/4) per axis (SplitMix64) and the mesh
re-triangulated (jittered_rectangle).  The reference's own small configuration
(50 x 50 plate, diffusion_equation_on_a_square_plate.jl) is in the tests.
"""
import numpy as np

from .conditions import BoundaryConditions, Dirichlet
from .forms import NodalDiffusion
from .geometry import FVMGeometry
from .mesh import jittered_rectangle
from .problem import FVMProblem

SEED = 20261202


def _zero(x, y, t, u, p):
    return 0.0 * u


def throughput_problem(n_side=2000, tile=0, seed=SEED):
    """(tri, mesh, prob, u): the RHS throughput workload."""
    tri = jittered_rectangle(0.0, 1.0, 0.0, 1.0, n_side, n_side, jitter=0.25, seed=seed)
    mesh = FVMGeometry(tri, tile_triangles=tile)
    BCs = BoundaryConditions(mesh, _zero, Dirichlet)
    u = np.random.default_rng(seed + 1).random(tri.num_points())
    prob = FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 0.1 + u ** 2,
                      source_function=lambda x, y, t, u, p: u * (1 - u),
                      initial_condition=u, final_time=1.0)
    return tri, mesh, prob, u


def template_D(x, y, p=None):
    """D(x, y) = exp(0.5 sin(4 pi x) cos(4 pi y)) of the steady template workload."""
    return np.exp(0.5 * np.sin(4 * np.pi * x) * np.cos(4 * np.pi * y))


def template_problems(n_side=1000, seed=SEED + 3):
    """(tri, poisson, mean_exit_time, f): the steady template workload."""
    from .templates import MeanExitTimeProblem, PoissonsEquation

    tri = jittered_rectangle(0.0, 1.0, 0.0, 1.0, n_side, n_side, jitter=0.25, seed=seed)
    mesh = FVMGeometry(tri)
    BCs = BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, Dirichlet)
    f = np.random.default_rng(seed + 1).uniform(-1.0, 1.0, tri.num_points())
    pois = PoissonsEquation(mesh, BCs, diffusion_function=template_D,
                            source_function=lambda x, y, p: p, source_parameters=f)
    met = MeanExitTimeProblem(mesh, BCs, diffusion_function=template_D)
    return tri, pois, met, f


def differentiation_problem(n_side=1000, seed=SEED + 7, final_time=0.1):
    """(tri, prob, u, logd, w): the per-node log-diffusivity workload."""
    tri = jittered_rectangle(0.0, 1.0, 0.0, 1.0, n_side, n_side, jitter=0.25, seed=seed)
    mesh = FVMGeometry(tri)
    rng = np.random.default_rng(seed + 1)
    N = tri.num_points()
    logd = np.log(0.1) + 0.3 * rng.standard_normal(N)
    u = rng.random(N)
    w = rng.standard_normal(N)
    BCs = BoundaryConditions(mesh, _zero, Dirichlet)
    prob = FVMProblem(mesh, BCs, diffusion_function=NodalDiffusion(logd),
                      initial_condition=u, final_time=final_time)
    return tri, prob, u, logd, w


def ensemble_sizes(P=256, seed=SEED + 5):
    """(nx, ny, D) of each ensemble member, drawn in member order."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(P):
        nx, ny = (int(round(np.exp(rng.uniform(np.log(20), np.log(300))))) for _ in range(2))
        D = float(np.exp(np.log(0.1) + 0.5 * rng.standard_normal()))
        out.append((nx, ny, D))
    return out


def ensemble_members(P=256, seed=SEED + 5, final_time=1.0):
    """(problems, D): the ensemble workload's members and their diffusivities."""
    probs, Ds = [], []
    for p, (nx, ny, D) in enumerate(ensemble_sizes(P, seed)):
        tri = jittered_rectangle(0.0, 2.0, 0.0, 2.0, nx, ny, jitter=0.25, seed=seed + 995 + p)
        mesh = FVMGeometry(tri)
        BCs = BoundaryConditions(mesh, _zero, Dirichlet)
        u0 = np.where(tri.y <= 1.0, 50.0, 0.0)
        probs.append(FVMProblem(mesh, BCs, diffusion_function=(lambda d: (lambda x, y, t, u, p: d + 0.0 * u))(D),
                                initial_condition=u0, final_time=final_time))
        Ds.append(D)
    return probs, Ds


def met_interface_radius(theta):
    """R1(theta) = 2 (1 + 0.05 (sin 3 theta + cos 5 theta)): the interface between the two
    diffusivities of the mean-exit-time walkthrough (mean_exit_time.jl:89-93)."""
    return 2.0 * (1.0 + 0.05 * (np.sin(3 * theta) + np.cos(5 * theta)))


def met_annulus_triangulation(h=0.1):
    """The mean-exit-time walkthrough's domain (mean_exit_time.jl:89-119), the annulus
    1 <= r <= 3, meshed here without DelaunayTriangulation's constrained refinement:
    the two point holes (-2, 0) and (0, 2.95) as points 0 and 1; 249 points on the
    interface r = R1(theta) (the reference's LinRange(0, 2 pi, 250) less its last);
    rings of points about h apart (ring points closer than h / 3 to those dropped), the
    outer ring holding the arc ends theta = +-0.25; Delaunay (Qhull), the triangles of
    the hole removed.  Boundary segments in the reference's order: 0 the outer arc
    0.25 <= theta <= 2 pi - 0.25, 1 the outer arc |theta| < 0.25, 2 the hole."""
    from scipy.spatial import cKDTree

    from .mesh import Triangulation, triangulate

    fixed = [(-2.0, 0.0), (0.0, 2.95)]
    th = np.linspace(0.0, 2 * np.pi, 250)[:-1]
    R = met_interface_radius(th)
    fixed += list(zip(R * np.cos(th), R * np.sin(th)))
    F = np.array(fixed)
    ring = []
    nr = int(round(2.0 / h))
    for k in range(nr + 1):
        r = 1.0 + 2.0 * k / nr
        if k == nr:
            m1 = max(2, int(round(0.5 * r / h)))
            m2 = max(4, int(round((2 * np.pi - 0.5) * r / h)))
            a = np.concatenate([np.linspace(-0.25, 0.25, m1 + 1), np.linspace(0.25, 2 * np.pi - 0.25, m2 + 1)[1:-1]])
        else:
            m = max(12, int(round(2 * np.pi * r / h)))
            a = 2 * np.pi * np.arange(m) / m + 0.37 * k
        ring += list(zip(r * np.cos(a), r * np.sin(a)))
    Rg = np.array(ring)
    rr = np.hypot(Rg[:, 0], Rg[:, 1])
    keep = (cKDTree(F).query(Rg)[0] >= h / 3) | (rr <= 1.0 + 1e-12) | (rr >= 3.0 - 1e-12)
    P = np.concatenate([F, Rg[keep]])
    full = triangulate(P[:, 0], P[:, 1])
    T = full.triangles
    T = T[np.hypot(P[T, 0].mean(1), P[T, 1].mean(1)) > 1.0]
    count = {}
    for a, b, c in T:
        for e in ((a, b), (b, c), (c, a)):
            count[e] = count.get(e, 0) + 1
    bedges = np.array(sorted(e for e in count if (e[1], e[0]) not in count), np.int32)
    mx, my = P[bedges, 0].mean(1), P[bedges, 1].mean(1)
    seg = np.where(np.hypot(mx, my) < 2.0, 2, np.where(np.abs(np.arctan2(my, mx)) < 0.25, 1, 0)).astype(np.int32)
    return Triangulation(P[:, 0], P[:, 1], T, bedges, seg)


def met_problems(h=0.1):
    """(tri, template, steady): the mean-exit-time walkthrough (mean_exit_time.jl:89-180)
    as the MeanExitTimeProblem template and as the SteadyFVMProblem of D grad^2 T = -1:
    D = 6.25e-4 inside the interface and 6.25e-5 outside, zero flux on the long outer
    arc, T = 0 on the short outer arc, the hole and the two point holes, the initial
    guess the radially symmetric T_exact (:163-170)."""
    from .conditions import InternalConditions, Neumann
    from .problem import SteadyFVMProblem
    from .templates import MeanExitTimeProblem

    tri = met_annulus_triangulation(h)
    mesh = FVMGeometry(tri)
    zero = lambda x, y, t, u, p: 0.0 * x  # noqa: E731
    BCs = BoundaryConditions(mesh, (zero, zero, zero), (Neumann, Dirichlet, Dirichlet))
    ICs = InternalConditions(zero, dirichlet_nodes={0: 0, 1: 0})
    D1, D2 = 6.25e-4, 6.25e-5

    def D(x, y, p=None):
        return np.where(np.hypot(x, y) < met_interface_radius(np.arctan2(y, x)), D1, D2)

    met = MeanExitTimeProblem(mesh, BCs, ICs, diffusion_function=D)
    r = np.hypot(tri.x, tri.y)
    R1, R2 = 2.0, 3.0
    guess = np.where(r < R1, (R1 ** 2 - r ** 2) / (4 * D1) + (R2 ** 2 - R1 ** 2) / (4 * D2), (R2 ** 2 - r ** 2) / (4 * D2))
    steady = SteadyFVMProblem(FVMProblem(mesh, BCs, ICs, diffusion_function=lambda x, y, t, u, p: D(x, y) + 0.0 * u,
                                         source_function=lambda x, y, t, u, p: 1.0 + 0.0 * u,
                                         initial_condition=guess, final_time=np.inf))
    return tri, met, steady
