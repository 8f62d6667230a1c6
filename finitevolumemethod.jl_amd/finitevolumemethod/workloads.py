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
