"""FVMSystem host logic (no GPU): tracing of the system closures onto the device
forms, the merged condition numbering (merge_problem_conditions, problem.jl:368-400),
the column-major state layout, the InvalidFluxError check (problem.jl:277-302)
and jacobian_sparsity for systems (solve.jl:96-131)."""
import numpy as np
import pytest

import finitevolumemethod as FVM
from finitevolumemethod import forms
from finitevolumemethod.system import system_physics
from oracle import geometry


class HostMesh:
    """The parts of FVMGeometry that set-up reads, from the oracle's geometry."""

    def __init__(self, tri):
        self.triangulation = tri
        self.cv_volumes, self.props_array = geometry(tri.x, tri.y, tri.triangles)


def brusselator(mesh, final_time=1.0):
    tri = mesh.triangulation
    phi = FVM.BoundaryConditions(mesh, (lambda x, y, t, u, p: -1 / 4 * np.exp(-x - t / 2),
                                        lambda x, y, t, u, p: 1 / 4 * np.exp(-1 - y - t / 2),
                                        lambda x, y, t, u, p: np.exp(-1 - x - t / 2),
                                        lambda x, y, t, u, p: -1 / 4 * np.exp(-y - t / 2)),
                                 (FVM.Neumann, FVM.Neumann, FVM.Dirichlet, FVM.Neumann))
    psi = FVM.BoundaryConditions(mesh, (lambda x, y, t, u, p: np.exp(x + t / 2),
                                        lambda x, y, t, u, p: -1 / 4 * np.exp(1 + y + t / 2),
                                        lambda x, y, t, u, p: -1 / 4 * np.exp(1 + x + t / 2),
                                        lambda x, y, t, u, p: np.exp(y + t / 2)),
                                 (FVM.Dirichlet, FVM.Neumann, FVM.Neumann, FVM.Dirichlet))
    p1 = FVM.FVMProblem(mesh, phi, flux_function=lambda x, y, t, a, b, g, p: (-a[0] / 4, -b[0] / 4),
                        source_function=lambda x, y, t, u, p: u[0] ** 2 * u[1] - 2 * u[0],
                        initial_condition=np.exp(-tri.x - tri.y), final_time=final_time)
    p2 = FVM.FVMProblem(mesh, psi, flux_function=lambda x, y, t, a, b, g, p: (-a[1] / 4, -b[1] / 4),
                        source_function=lambda x, y, t, u, p: -u[0] ** 2 * u[1] + u[0],
                        initial_condition=np.exp(tri.x + tri.y), final_time=final_time)
    return FVM.FVMSystem(p1, p2)


@pytest.fixture(scope="module")
def mesh():
    return HostMesh(FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 9, 8, single_boundary=False))


def test_brusselator_traces_onto_the_system_forms(mesh):
    sys = brusselator(mesh)
    tri = mesh.triangulation
    ph = system_physics(sys)
    assert np.array_equal(ph["flux_k"], [[0.25, 0.0], [0.0, 0.25]])
    assert sorted(ph["terms"]) == sorted([(0, (2, 1), 1.0), (0, (1, 0), -2.0),
                                          (1, (2, 1), -1.0), (1, (1, 0), 1.0)])
    assert ph["s_node"] is None and not ph["st"].any()
    # merged numbering: Phi's four functions, then Psi's
    assert sys.cnum_fncs == (0, 4) and len(sys.functions) == 8
    assert np.array_equal(ph["fn_lam"], [-0.5] * 4 + [0.5] * 4)
    assert not ph["fn_cu"].any() and not ph["fn_c0"].any()
    # Psi is Dirichlet on the bottom: a = exp(x) exp(t / 2), g = exp(x) at those nodes
    # (the corner (0, 0) takes the left segment's function, the later one: conditions.jl)
    bottom = np.nonzero(ph["node_fidx"][1] == 4)[0]
    assert np.array_equal(bottom, np.nonzero((tri.y == 0.0) & (tri.x > 0.0))[0])
    assert np.all(ph["node_kind"][1, bottom] == FVM._lib.NODE_DIRICHLET)
    assert ph["node_fidx"][1, np.nonzero((tri.x == 0.0) & (tri.y == 0.0))[0][0]] == 7
    np.testing.assert_allclose(ph["node_g"][1, bottom], np.exp(tri.x[bottom]), rtol=1e-15)
    # Phi's bottom edges are Neumann with function 0; g at the half-edge points
    be = tri.boundary_edges
    e0 = np.nonzero(ph["bedge_neumann_fidx"][0] == 0)[0]
    assert len(e0) == 8 and np.all(tri.y[be[e0]] == 0.0)
    mx = (3 * tri.x[be[e0, 0]] + tri.x[be[e0, 1]]) / 4
    np.testing.assert_allclose(ph["bedge_g"][0, 2 * e0], -0.25 * np.exp(-mx), rtol=1e-14)


def test_general_condition_and_source_forms():
    x = np.array([0.1, 0.4, 0.9])
    y = np.array([0.3, 0.2, 0.7])
    r = forms.classify_system_condition(lambda x, y, t, u: 2 * u[1] - 0.5 * u[0] + x * y + 3 * t, 2, x, y)
    assert r["cu"] == [-0.5, 2.0] and r["ct"] == 3.0 and r["lam"] == 0.0 and r["c0"] == 0.0
    np.testing.assert_array_equal(r["g"], x * y)
    r = forms.classify_system_condition(lambda x, y, t, u: 2.0 + np.exp(x + 0.5 * t), 2, x, y)
    assert r["c0"] == 2.0 and r["lam"] == 0.5
    np.testing.assert_allclose(r["g"], np.exp(x), rtol=1e-15)
    sn, terms, st = forms.classify_system_source(lambda x, y, t, u: 1 + x * y - u[0] * u[1] ** 2 + 0.5 * t,
                                                 2, x, y)
    np.testing.assert_allclose(sn, 1 + x * y)
    assert terms == [((1, 2), -1.0)] and st == 0.5
    K = forms.classify_system_flux(lambda x, y, t, a, b, g: (-(a[0] + 0.2 * a[1]), -(b[0] + 0.2 * b[1])), 2)
    assert K == [1.0, 0.2]


@pytest.mark.parametrize("fn,kind", [
    (lambda x, y, t, a, b, g: (-a[0] * (a[0] * x + g[0]), -b[0]), "flux"),   # u-dependent D
    (lambda x, y, t, a, b, g: (-a[0] * x, -b[0] * x), "flux"),               # D(x, y)
    (lambda x, y, t, a, b, g: (-a[0], -2 * b[0]), "flux"),                   # anisotropic
    (lambda x, y, t, a, b, g: (-a[0] * np.exp(-t), -b[0] * np.exp(-t)), "flux"),
    (lambda x, y, t, u: u[0] ** 4, "source"),
    (lambda x, y, t, u: np.exp(u[0]), "source"),
    (lambda x, y, t, u: x * u[0], "source"),
    (lambda x, y, t, u: u[0] * u[1], "cond"),
    (lambda x, y, t, u: x * np.exp(-t) + y, "cond"),
    (lambda x, y, t, u: np.exp(-t) + np.exp(-2 * t), "cond"),
    (lambda x, y, t, u: np.exp(-x * t), "cond"),
    (lambda x, y, t, u: u[0] if t > 0 else 0.0, "cond"),
])
def test_untraceable_closures_raise(fn, kind):
    x = np.array([0.1, 0.4, 0.9])
    y = np.array([0.3, 0.2, 0.7])
    with pytest.raises(forms.UnsupportedClosure):
        if kind == "flux":
            forms.classify_system_flux(fn, 2)
        elif kind == "source":
            forms.classify_system_source(fn, 2, x, y)
        else:
            forms.classify_system_condition(fn, 2, x, y)


def test_scalar_conditions_trace_exponentials_in_time():
    """exp(lam t) conditions trace for scalar problems too; two rates do not."""
    r = forms.classify_condition(lambda x, y, t, u: np.exp(-t) + 0.0 * u, np.array([0.5]), np.array([0.5]))
    assert r["lam"] == -1.0 and r["g"][0] == 1.0 and r["c0"] == 0.0
    with pytest.raises(forms.UnsupportedClosure):
        forms.classify_condition(lambda x, y, t, u: np.exp(-t) + np.exp(-2 * t), np.array([0.5]), np.array([0.5]))


def test_diffusion_only_problem_is_an_invalid_flux(mesh):
    tri = mesh.triangulation
    BCs = FVM.BoundaryConditions(mesh, (lambda x, y, t, u, p: 0.0,) * 4, (FVM.Dirichlet,) * 4)
    p1 = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1.0,
                        initial_condition=np.zeros(tri.num_points()), final_time=1.0)
    with pytest.raises(FVM.InvalidFluxError):
        FVM.FVMSystem(p1, p1)


def test_problems_must_agree(mesh):
    sys = brusselator(mesh)
    p1, p2 = sys.problems
    BCs = FVM.BoundaryConditions(mesh, (lambda x, y, t, u, p: 0.0,) * 4, (FVM.Neumann,) * 4)
    p3 = FVM.FVMProblem(mesh, BCs, flux_function=lambda x, y, t, a, b, g, p: (-a[1], -b[1]),
                        initial_condition=p2.initial_condition, final_time=2.0)
    with pytest.raises(AssertionError):
        FVM.FVMSystem(p1, p3)
    with pytest.raises(ValueError):
        FVM.FVMSystem()


def test_state_layout_is_column_major(mesh):
    sys = brusselator(mesh)
    U = sys.initial_condition
    f = sys.flat(U)
    N = U.shape[1]
    assert f[0] == U[0, 0] and f[1] == U[1, 0] and f[2 * 5 + 1] == U[1, 5]
    assert np.array_equal(sys.unflat(f), U)
    assert f.shape == (2 * N,) and sys.n_vars == 2 and FVM.is_system(sys)
    assert FVM.is_system(FVM.SteadyFVMProblem(sys)) and not FVM.is_system(sys.problems[0])
    assert repr(sys) == "FVMSystem with 2 equations and time span (0.0, 1.0)"


def test_system_jacobian_sparsity_follows_the_reference_construction(mesh):
    """solve.jl:96-131 built literally: (i, l) related to (j, k) for all variables when
    j = i or j neighbours i."""
    sys = brusselator(mesh)
    tri = mesh.triangulation
    nv, N = 2, tri.num_points()
    A = FVM.jacobian_sparsity(sys)
    nbrs = [set() for _ in range(N)]
    for a, b, c in tri.triangles:
        for p, q in ((a, b), (b, c), (c, a)):
            nbrs[p].add(q)
            nbrs[q].add(p)
    want = np.zeros((nv * N, nv * N), bool)
    for i in range(N):
        for j in [i] + sorted(nbrs[i]):
            for l_ in range(nv):
                for k in range(nv):
                    want[i * nv + l_, j * nv + k] = True
    assert np.array_equal(A.toarray() != 0, want)
