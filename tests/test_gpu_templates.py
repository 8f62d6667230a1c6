"""Linear templates on the GPU (src/specific_problems/*.jl): device assembly
against the oracle's restatement of abstract_templates.jl:73-325, the device
solvers against a sparse direct solve of the oracle's matrix and against the
closed forms the reference's docs test (poissons_equation.jl:154-162,
laplaces_equation.jl:159-193, the disk mean exit time), and the ODE templates
against the general FVMProblem path."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import finitevolumemethod as FVM
from finitevolumemethod import templates as TM
from problems import disk_triangulation, template_oracle

pytestmark = pytest.mark.gpu

Dxy = lambda x, y, p: np.exp(0.5 * np.sin(4 * np.pi * x) * np.cos(4 * np.pi * y))  # noqa: E731


def _oracle_matrix(prob, kind, D, **kw):
    orc = template_oracle(prob)
    r, c, v, b = orc.template_assemble(kind, D, **kw)
    n = prob.problem.n
    A = sp.csr_matrix((v, (r, c)), shape=(n, n))
    A.sum_duplicates()
    absA = sp.csr_matrix((np.abs(v), (r, c)), shape=(n, n))
    return A, b, np.asarray(absA.sum(axis=1)).ravel()


def _assert_same_operator(prob, kind, D, **kw):
    A_ref, b_ref, rowscale = _oracle_matrix(prob, kind, D, **kw)
    A = prob.A
    n = A.shape[0]
    # the device pattern is jacobian_sparsity's, bit-exact
    rp, col = FVM.equations.csr_pattern(prob.mesh.triangulation)
    np.testing.assert_array_equal(A.indptr, rp)
    np.testing.assert_array_equal(A.indices, col)
    # every oracle entry lies in the pattern, and the values agree to rounding
    D_ = (A - A_ref).tocoo()
    scale = np.maximum(rowscale, 1e-300)
    assert np.all(np.abs(D_.data) <= 1e-13 * scale[D_.row]), np.max(np.abs(D_.data) / scale[D_.row])
    pattern = sp.csr_matrix((np.ones(A.nnz), A.indices, A.indptr), shape=A.shape)
    ref = A_ref.tocoo()
    assert np.all(np.asarray(pattern[ref.row, ref.col]).ravel() == 1.0)
    bs = np.max(np.abs(b_ref)) if n else 1.0
    np.testing.assert_allclose(prob.b, b_ref, rtol=1e-13, atol=1e-15 * bs)
    return A_ref, b_ref


def _rect4(types, fns, nx=23, ny=19, jitter=0.2, seed=8, params=None):
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, nx, ny, jitter=jitter, seed=seed,
                                 single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    return tri, mesh, FVM.BoundaryConditions(mesh, fns, types, parameters=params)


def test_poisson_assembly_matches_oracle(gpu):
    types = (FVM.Neumann, FVM.Constrained, FVM.Dirichlet, FVM.Neumann)
    fns = (lambda x, y, t, u, p: p + x, lambda x, y, t, u, p: 0.0,
           lambda x, y, t, u, p: 1 + x * y, lambda x, y, t, u, p: -0.5 * y)
    tri, mesh, BCs = _rect4(types, fns, params=(0.3, None, None, None))
    nx = 23
    ICs = FVM.InternalConditions(lambda x, y, t, u, p: x - y, dirichlet_nodes={5 + 7 * nx: 0, 9 + 11 * nx: 0})
    src = lambda x, y, p: x ** 2 - p * y  # noqa: E731
    prob = FVM.PoissonsEquation(mesh, BCs, ICs, diffusion_function=Dxy, source_function=src,
                                source_parameters=2.0)
    kind, fidx, bn = prob.conditions.flatten(tri)
    mx, my = FVM.problem._half_edge_points(tri)
    neu_a = np.zeros(2 * len(tri.boundary_edges))
    for e in np.nonzero(bn >= 0)[0]:
        f = prob.conditions.functions[bn[e]]
        neu_a[2 * e:2 * e + 2] = [f(mx[2 * e + k], my[2 * e + k], None, None) for k in (0, 1)]
    dir_val = np.array([prob.conditions.functions[fidx[i]](tri.x[i], tri.y[i], None, None)
                        if kind[i] == 1 else 0.0 for i in range(tri.num_points())])
    _assert_same_operator(prob, 0, lambda x, y: float(Dxy(x, y, None)),
                          f_node=src(tri.x, tri.y, 2.0), neu_a=neu_a, dir_val=dir_val)


def test_mean_exit_time_assembly_matches_oracle(gpu):
    types = (FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet, FVM.Neumann)
    fns = tuple(lambda x, y, t, u, p: 0.0 for _ in range(4))
    tri, mesh, BCs = _rect4(types, fns, nx=21, ny=24, seed=2)
    D = lambda x, y, p: 1 + x + y ** 2  # noqa: E731
    prob = FVM.MeanExitTimeProblem(mesh, BCs, diffusion_function=D)
    _assert_same_operator(prob, 1, lambda x, y: 1 + x + y ** 2)


@pytest.mark.parametrize("reaction", [False, True])
def test_diffusion_assembly_matches_oracle(gpu, reaction):
    types = (FVM.Neumann, FVM.Dudt, FVM.Dirichlet, FVM.Constrained)
    fns = (lambda x, y, t, u, p: 0.3 + x, lambda x, y, t, u, p: 1.5 - y,
           lambda x, y, t, u, p: 0.0, lambda x, y, t, u, p: 0.0)
    tri, mesh, BCs = _rect4(types, fns, seed=11)
    u0 = np.random.default_rng(0).random(tri.num_points())
    kw = dict(diffusion_function=Dxy, initial_condition=u0, final_time=0.1)
    if reaction:
        prob = FVM.LinearReactionDiffusionEquation(mesh, BCs, source_function=lambda x, y, p: 0.3 - x, **kw)
    else:
        prob = FVM.DiffusionEquation(mesh, BCs, **kw)
    kind, fidx, bn = prob.conditions.flatten(tri)
    mx, my = FVM.problem._half_edge_points(tri)
    neu_a = np.zeros(2 * len(tri.boundary_edges))
    for e in np.nonzero(bn >= 0)[0]:
        neu_a[2 * e:2 * e + 2] = 0.3 + mx[2 * e:2 * e + 2]
    dudt = np.where(kind == 2, 1.5 - tri.y, 0.0)
    A_ref, b_ref = _assert_same_operator(prob, 2, lambda x, y: float(Dxy(x, y, None)), neu_a=neu_a,
                                         dudt_val=dudt, r_node=(0.3 - tri.x) if reaction else None)
    # du = A u + b on the device
    u = FVM.DeviceArray.from_host(u0)
    du = FVM.DeviceArray(len(u0))
    prob.problem.apply(u, du)
    ref = A_ref @ u0 + b_ref
    scale = abs(A_ref) @ np.abs(u0) + np.abs(b_ref)
    assert np.all(np.abs(du.to_host() - ref) <= 1e-13 * scale)


def test_points_that_are_not_vertices_get_identity_rows(gpu):
    base = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 15, 13, jitter=0.25, seed=4)
    x = np.concatenate([base.x, [0.5, 0.25]])
    y = np.concatenate([base.y, [0.5, 0.75]])
    tri = FVM.Triangulation(x, y, base.triangles, base.boundary_edges, base.boundary_segments)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    n = len(x)
    # an internal Dirichlet condition on a point that is not a vertex keeps A_ii = 1
    ICs = FVM.InternalConditions(lambda x, y, t, u, p: 2.0, dirichlet_nodes={n - 1: 0})
    prob = FVM.PoissonsEquation(mesh, BCs, ICs, source_function=lambda x, y, p: 1.0)
    kind, fidx, _ = prob.conditions.flatten(tri)
    f = np.where(kind == 1, 0.0, 1.0)
    f[-2:] = 1.0
    dv = np.where(kind == 1, 0.0, 0.0)
    dv[-1] = 2.0
    _assert_same_operator(prob, 0, 1.0, f_node=f, dir_val=dv)
    A = prob.A
    assert A[n - 1, n - 1] == 1.0 and A[n - 2, n - 2] == 1.0
    assert prob.b[-1] == 0.0 and prob.b[-2] == 0.0
    sol = FVM.solve(prob, TM.CG())
    assert sol.retcode == "Success" and sol.u[-1] == 0.0 and sol.u[-2] == 0.0


ALGS = [TM.CG(), TM.MixedCG(), TM.MixedCG(update="restart"), TM.BiCGStab()]
IDS = ["cg", "mixed", "mixed_restart", "bicgstab"]


@pytest.mark.parametrize("alg", ALGS, ids=IDS)
def test_poisson_closed_form(gpu, alg):
    """poissons_equation.jl:154-162: laplace(u) = -sin(pi x) sin(pi y) on [0,1]^2
    (100 x 100), u = 0 on the boundary: u = sin(pi x) sin(pi y) / (2 pi^2), rtol 1e-4."""
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 100, 100, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    prob = FVM.PoissonsEquation(mesh, BCs, source_function=lambda x, y, p: -np.sin(np.pi * x) * np.sin(np.pi * y))
    sol = FVM.solve(prob, alg)
    assert sol.retcode == "Success", sol.stats
    exact = np.sin(np.pi * tri.x) * np.sin(np.pi * tri.y) / (2 * np.pi ** 2)
    assert np.linalg.norm(sol.u - exact) <= 1e-4 * np.linalg.norm(exact)
    direct = spla.spsolve(prob.A.tocsc(), prob.b)
    assert np.linalg.norm(sol.u - direct) <= 1e-6 * np.linalg.norm(direct)


@pytest.mark.parametrize("alg", ALGS, ids=IDS)
def test_laplace_closed_form(gpu, alg):
    """laplaces_equation.jl:159-193: D = (x+1)(y+2) on [0,5]^2, Neumann bottom / top,
    u = 5 right, u = 0 left: u = 5 log(1 + x) / log 6, rtol 1e-3."""
    tri = FVM.triangulate_rectangle(0.0, 5.0, 0.0, 5.0, 100, 100, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    zero = lambda x, y, t, u, p: 0.0  # noqa: E731
    five = lambda x, y, t, u, p: 5.0  # noqa: E731
    BCs = FVM.BoundaryConditions(mesh, (zero, five, zero, zero),
                                 (FVM.Neumann, FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet))
    prob = FVM.LaplacesEquation(mesh, BCs, diffusion_function=lambda x, y, p: (x + 1) * (y + 2))
    sol = FVM.solve(prob, alg)
    assert sol.retcode == "Success", sol.stats
    exact = 5 * np.log(1 + tri.x) / np.log(6)
    assert np.linalg.norm(sol.u - exact) <= 1e-3 * np.linalg.norm(exact)


def test_mean_exit_time_disk_closed_form(gpu):
    """T(r) = (R^2 - r^2) / (4 D) on a disk with an absorbing boundary."""
    tri = disk_triangulation(nr=24, R=2.0)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    prob = FVM.MeanExitTimeProblem(mesh, BCs, diffusion_function=lambda x, y, p: 0.3)
    _assert_same_operator(prob, 1, 0.3)
    sol = FVM.solve(prob, TM.CG())
    exact = (4.0 - (tri.x ** 2 + tri.y ** 2)) / (4 * 0.3)
    assert np.linalg.norm(sol.u - exact) <= 5e-3 * np.linalg.norm(exact)


@pytest.mark.parametrize("alg", ALGS, ids=IDS)
def test_solvers_reach_the_true_residual(gpu, alg):
    """The variable-D Poisson system at 200 x 200 (jittered): every method stops at
    |b - A x| <= 1e-10 |b|, checked on the host with the oracle's own matrix."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 200, 200, jitter=0.25, seed=20261202)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    f = np.random.default_rng(7).uniform(-1, 1, tri.num_points())
    prob = FVM.PoissonsEquation(mesh, BCs, diffusion_function=Dxy,
                                source_function=lambda x, y, p: p, source_parameters=f)
    kind, _, _ = prob.conditions.flatten(tri)
    A_ref, b_ref, _ = _oracle_matrix(prob, 0, "exp_sincos:1.0", f_node=np.where(kind == 1, 0.0, f))
    sol = FVM.solve(prob, alg, rtol=1e-10)
    assert sol.retcode == "Success", sol.stats
    rel = np.linalg.norm(b_ref - A_ref @ sol.u) / np.linalg.norm(b_ref)
    assert rel <= 1.5e-10, (rel, sol.stats)
    assert abs(sol.resid - rel) <= 0.05 * rel + 1e-14
    direct = spla.spsolve(A_ref.tocsc(), b_ref)
    assert np.linalg.norm(sol.u - direct) <= 1e-5 * np.linalg.norm(direct)


def test_solve_is_bitwise_reproducible(gpu):
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 60, 60, jitter=0.25, seed=3)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    prob = FVM.PoissonsEquation(mesh, BCs, diffusion_function=Dxy, source_function=lambda x, y, p: x - y)
    for alg in ALGS:
        a = FVM.solve(prob, alg).u
        b = FVM.solve(prob, alg).u
        assert np.array_equal(a, b)


@pytest.mark.parametrize("alg", [TM.CG, lambda **k: TM.MixedCG(update="restart", **k), TM.BiCGStab],
                         ids=["cg", "mixed_restart", "bicgstab"])
def test_iteration_count_does_not_depend_on_the_check_interval(gpu, alg):
    """The host reads the solver scalars every check_every iterations; launches past
    the stopping iteration are no-ops, so reading after every iteration and every 32
    gives the same iteration count and the same solution bit for bit."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 40, 40, jitter=0.25, seed=4)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    prob = FVM.PoissonsEquation(mesh, BCs, diffusion_function=Dxy, source_function=lambda x, y, p: x - y)
    one = FVM.solve(prob, alg(check_every=1))
    many = FVM.solve(prob, alg(check_every=32))
    assert one.retcode == "Success" and one.iters == many.iters, (one.stats, many.stats)
    assert np.array_equal(one.u, many.u)


def test_fp32_operator_keeps_accuracy_on_smooth_vectors(gpu):
    """The FP32 solver operator (difference form) against the FP64 one, on vectors
    that FP32 holds exactly (inside the solver every vector is FP32): on a smooth
    vector, whose product cancels to about 1e-4 of |M| |p|, the FP32 product stays
    within 1e-4 of |M p| (the direct FP32 form loses about 1e-3 to cancellation);
    on a rough vector within 1e-6."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 300, 300, jitter=0.25, seed=6)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    dt = FVM.MeanExitTimeProblem(mesh, BCs, diffusion_function=Dxy).problem
    rp, col, val, _ = dt.download()
    n = dt.n
    free = dt.node_kind == 0
    absA = sp.csr_matrix((np.abs(val), col, rp), shape=(n, n))
    for x, cancel, tol in ((np.sin(np.pi * tri.x) * np.sin(np.pi * tri.y) * free, True, 1e-4),
                           (np.random.default_rng(1).uniform(-1, 1, n) * free, False, 1e-6)):
        x = x.astype(np.float32).astype(np.float64)
        d_x = FVM.DeviceArray.from_host(x)
        y64, y32 = FVM.DeviceArray(n), FVM.DeviceArray(n)
        dt.spmv(1, d_x, y64)
        dt.spmv(2, d_x, y32)
        a, b = y64.to_host(), y32.to_host()
        A = sp.csr_matrix((val, col, rp), shape=(n, n))
        assert (np.linalg.norm(A @ x) < 1e-3 * np.linalg.norm(absA @ np.abs(x))) == cancel
        assert np.linalg.norm(b - a) <= tol * np.linalg.norm(a)


def test_reliable_updates_keep_the_iteration_count_of_fp64_cg(gpu):
    """One FP32 CG sequence with reliable updates needs about as many iterations as
    FP64 CG (within 15 %) and fewer than restarted refinement, and reaches the same
    true residual; its solution agrees with FP64 CG to the tolerance."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 150, 150, jitter=0.25, seed=9)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    for prob in (FVM.PoissonsEquation(mesh, BCs, diffusion_function=Dxy, source_function=lambda x, y, p: x - y),
                 FVM.MeanExitTimeProblem(mesh, BCs, diffusion_function=Dxy)):
        cg = FVM.solve(prob, TM.CG(), rtol=1e-10)
        rel = FVM.solve(prob, TM.MixedCG(), rtol=1e-10)
        rst = FVM.solve(prob, TM.MixedCG(update="restart"), rtol=1e-10)
        assert cg.retcode == rel.retcode == rst.retcode == "Success"
        assert rel.resid <= 1e-10 and rel.stats["outer"] >= 2
        assert rel.iters <= 1.15 * cg.iters, (rel.stats, cg.stats)
        assert rel.iters < rst.iters, (rel.stats, rst.stats)
        assert np.linalg.norm(rel.u - cg.u) <= 1e-8 * np.linalg.norm(cg.u)


def test_diffusion_template_matches_fvm_problem(gpu):
    """The DiffusionEquation template and the general FVMProblem give the same
    solution (diffusion_equations.jl:345-347 compares them the same way)."""
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, 30, 30, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    u0 = np.where(tri.y <= 1.0, 50.0, 0.0)
    u0[tri.boundary_nodes()] = 0.0  # the template applies the Dirichlet values to u0 itself
    tmpl = FVM.DiffusionEquation(mesh, BCs, diffusion_function=lambda x, y, p: 1 / 9,
                                 initial_condition=u0, final_time=0.5)
    gen = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1 / 9,
                         initial_condition=u0, final_time=0.5)
    s1 = FVM.solve(tmpl, FVM.Tsit5(), saveat=0.1, abstol=1e-10, reltol=1e-10)
    s2 = FVM.solve(gen, FVM.Tsit5(), saveat=0.1, abstol=1e-10, reltol=1e-10)
    assert len(s1.u) == len(s2.u) == 6
    for a, b in zip(s1.u, s2.u):
        assert len(a) == tri.num_points() + 1 and a[-1] == 1.0
        assert np.max(np.abs(a[:-1] - b)) <= 1e-7 * 50


def test_template_argument_errors(gpu):
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 10, 10, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    z = lambda x, y, t, u, p: 0.0  # noqa: E731
    BCs = FVM.BoundaryConditions(mesh, (z, z, z, z), (FVM.Dirichlet, FVM.Dudt, FVM.Neumann, FVM.Constrained))
    with pytest.raises(ValueError):
        FVM.PoissonsEquation(mesh, BCs, source_function=lambda x, y, p: 1.0)
    with pytest.raises(ValueError):
        FVM.MeanExitTimeProblem(mesh, BCs, diffusion_function=lambda x, y, p: 1.0)
    BCs2 = FVM.BoundaryConditions(mesh, (z, z, z, z), (FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet, FVM.Constrained))
    with pytest.raises(ValueError):
        FVM.MeanExitTimeProblem(mesh, BCs2, diffusion_function=lambda x, y, p: 1.0)
    d = FVM.DiffusionEquation(mesh, BCs2, diffusion_function=lambda x, y, p: 1.0,
                              initial_condition=np.zeros(tri.num_points()), final_time=1.0)
    with pytest.raises(FVM.FVMError):
        d.problem.solve(TM.CG())
    p = FVM.PoissonsEquation(mesh, BCs2, source_function=lambda x, y, p: 1.0)
    with pytest.raises(TypeError):
        p.problem.solve(object())


def test_large_diffusion_template_matches_fvm_problem(gpu):
    """diffusion_equations.jl:343-452: the 100 x 100 DiffusionEquation on [0, 320]^2
    (D = 2, Neumann 2 in the template's convention, -4 as the FVMProblem's q . n) and
    the equivalent FVMProblem, both with Tsit5 to t = 500, saveat 100: the reference
    asserts rtol 1e-3 between them; here they agree to 1e-8."""
    L = 320.0
    tri = FVM.triangulate_rectangle(0.0, L, 0.0, L, 100, 100, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    u0 = np.where((0.4 * L <= tri.y) & (tri.y <= 0.6 * L), 1.0, 0.0)
    tm = FVM.DiffusionEquation(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 2.0, FVM.Neumann),
                               diffusion_function=lambda x, y, p: 2.0, initial_condition=u0, final_time=500.0)
    fp = FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: -4.0 + 0.0 * u, FVM.Neumann),
                        diffusion_function=lambda x, y, t, u, p: 2.0 + 0.0 * u, initial_condition=u0,
                        final_time=500.0)
    a = FVM.solve(tm, FVM.Tsit5(), saveat=100.0)
    b = FVM.solve(fp, FVM.Tsit5(), saveat=100.0)
    assert a.retcode == b.retcode == "Success" and len(a.u) == len(b.u) == 6
    ua = np.array([u[:-1] for u in a.u[1:]])
    ub = np.array(b.u[1:])
    assert np.all(np.array([u[-1] for u in a.u]) == 1.0)
    assert np.max(np.abs(ua - ub)) <= 1e-8 * np.max(np.abs(ub))


def wheel_triangulation(k=40):
    """A centre point inside a ring of k points (its row has k off-diagonals: the
    ELL copy keeps the slots beyond its width in the overflow list), then rings of
    40, 60 and 80 points out to radius 2."""
    pts = [(0.0, 0.0)]
    for m, r in ((k, 0.5), (40, 1.0), (60, 1.5), (80, 2.0)):
        th = 2 * np.pi * np.arange(m) / m + 0.05 * r
        pts += list(zip(r * np.cos(th), r * np.sin(th)))
    P = np.array(pts)
    return FVM.triangulate(P[:, 0], P[:, 1])


@pytest.mark.parametrize("mesh", ["jittered", "wheel"])
@pytest.mark.parametrize("alg", ALGS, ids=IDS)
def test_ell_and_csr_forms_of_the_krylov_operator_agree(gpu, mesh, alg):
    """The Krylov products on the sliced ELL copy of M (a wave per 64-row slice, rows
    longer than the unrolled slots continued slot by slot) and on CSR: both
    converge to the direct solution of the same
    A u = b (1e-8 relative), and their iteration counts agree within a few percent
    (the two forms round the products' sums differently)."""
    if mesh == "wheel":
        tri = wheel_triangulation()
        rp, _ = FVM.equations.csr_pattern(tri)
        assert np.diff(rp).max() - 1 == 40
    else:
        tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 90, 80, jitter=0.25, seed=21)
    m = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(m, lambda x, y, t, u, p: 0.0 * x, FVM.Dirichlet)
    prob = FVM.PoissonsEquation(m, BCs, diffusion_function=lambda x, y, p: 1.0 + 0.5 * np.sin(3 * x) ** 2,
                                source_function=lambda x, y, p: np.cos(2 * x) - y)
    W, n_over, ib = prob.problem.ell_info(index_bytes=True)
    assert 4 <= W <= 10
    assert (n_over == 40 - W) if mesh == "wheel" else (W == 8 and n_over == 0)
    assert ib == 2  # every column within 2^15 of its row: 16-bit offsets
    direct = spla.spsolve(prob.A.tocsc(), prob.b)
    sols = {f: FVM.solve(prob, alg, rtol=1e-11, spmv_format=f) for f in ("ell", "csr")}
    for f, s in sols.items():
        assert s.retcode == "Success", (f, s.stats)
        assert np.linalg.norm(s.u - direct) <= 1e-8 * np.linalg.norm(direct), f
    it_e, it_c = sols["ell"].stats["iterations"], sols["csr"].stats["iterations"]
    assert abs(it_e - it_c) <= 0.05 * it_c + 3, (it_e, it_c)
    # the default form is the ELL one, and both are deterministic
    again = FVM.solve(prob, alg, rtol=1e-11)
    assert np.array_equal(again.u, sols["ell"].u)


def _permuted(base, seed):
    """The same triangulation with its nodes numbered in a random order."""
    perm = np.random.default_rng(seed).permutation(base.num_points())  # new id of each old node
    inv = np.argsort(perm)
    return FVM.Triangulation(base.x[inv], base.y[inv], perm[base.triangles], perm[base.boundary_edges],
                             base.boundary_segments)


@pytest.mark.parametrize("mesh", ["jittered", "wheel", "permuted"])
def test_sliced_ell_product_sums_each_row_in_csr_order(gpu, mesh):
    """y = A x on the sliced ELL copy (rows ordered by length inside each 64-row
    slice, no padding, rows longer than the unrolled slots continued slot by slot;
    16-bit column offsets, or 32-bit columns for the permuted numbering) adds each
    row's off-diagonal terms in CSR order and the diagonal term last: bit for bit
    the sum formed here in that order.  Meshes: 7,200 nodes (a partial last
    slice), a wheel whose hub has 40 neighbours, 38,000 nodes randomly renumbered."""
    if mesh == "wheel":
        tri = wheel_triangulation()
    elif mesh == "permuted":
        tri = _permuted(FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 200, 190, jitter=0.25, seed=3), 3)
    else:
        tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 90, 80, jitter=0.25, seed=21)
    m = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(m, lambda x, y, t, u, p: 0.0 * x, FVM.Dirichlet)
    prob = FVM.PoissonsEquation(m, BCs, diffusion_function=lambda x, y, p: 1.0 + 0.5 * np.sin(3 * x) ** 2,
                                source_function=lambda x, y, p: np.cos(2 * x) - y)
    W, n_over, ib = prob.problem.ell_info(index_bytes=True)
    assert ib == (4 if mesh == "permuted" else 2)
    rp, col, val, _ = prob.problem.download()
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    on_diag = col == rows
    assert np.array_equal(np.bincount(rows[on_diag], minlength=n), np.ones(n))
    pos = np.arange(len(col)) - rp[rows]
    dpos = np.zeros(n, np.int64)
    dpos[rows[on_diag]] = pos[on_diag]
    slot = pos - (pos > dpos[rows])  # CSR position among the row's off-diagonals
    deg = np.bincount(rows[~on_diag], minlength=n)
    assert W == min(10, max(4, deg.max())) and n_over == np.maximum(deg - W, 0).sum()
    x = np.random.default_rng(5).uniform(-1.0, 1.0, n)
    acc = np.zeros(n)
    for s in range(deg.max()):
        sel = ~on_diag & (slot == s)
        term = np.zeros(n)
        term[rows[sel]] = val[sel] * x[col[sel]]
        acc = acc + term
    acc = acc + val[on_diag][np.argsort(rows[on_diag])] * x
    y = FVM.DeviceArray(n)
    prob.problem.spmv(0, FVM.DeviceArray.from_host(x), y)
    assert np.array_equal(y.to_host(), acc)


def test_unknown_spmv_format_is_rejected(gpu):
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 8, 8, single_boundary=True)
    m = FVM.FVMGeometry(tri)
    prob = FVM.PoissonsEquation(m, FVM.BoundaryConditions(m, lambda x, y, t, u, p: 0.0 * x, FVM.Dirichlet),
                                source_function=lambda x, y, p: 1.0)
    with pytest.raises(KeyError):
        FVM.solve(prob, spmv_format="dense")


def test_template_solve_forwards_callback_and_options(gpu):
    """solve(::AbstractFVMTemplate, alg; kwargs...) hands every keyword to the ODEProblem's
    solve (abstract_templates.jl:58-60): a callback runs once per accepted step, maxiters
    and save_start apply; a steady template rejects options a LinearProblem has no use for."""
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, 20, 20, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet)
    prob = FVM.DiffusionEquation(mesh, BCs, diffusion_function=lambda x, y, p: 1 / 9,
                                 initial_condition=np.where(tri.y <= 1.0, 50.0, 0.0), final_time=0.1)
    calls = []
    sol = FVM.solve(prob, FVM.Tsit5(), saveat=0.05, callback=lambda u, t: calls.append((t, u.to_host()[0])))
    assert sol.retcode == "Success"
    assert len(calls) == sol.stats["naccept"] > 3
    ts = [c[0] for c in calls]
    assert all(b > a for a, b in zip(ts, ts[1:])) and abs(ts[-1] - 0.1) <= 1e-12
    plain = FVM.solve(prob, FVM.Tsit5(), saveat=0.05)
    assert len(plain.u) == len(sol.u) == 3
    assert np.max(np.abs(plain.u[-1] - sol.u[-1])) <= 1e-12 * np.max(np.abs(plain.u[-1]))
    short = FVM.solve(prob, FVM.Tsit5(), saveat=0.05, maxiters=2, callback=lambda u, t: None)
    assert short.retcode == "MaxIters" and short.stats["naccept"] + short.stats["nreject"] == 2
    nostart = FVM.solve(prob, FVM.Tsit5(), saveat=0.05, save_start=False, callback=lambda u, t: None)
    assert nostart.t[0] > 0.0
    pois = FVM.PoissonsEquation(mesh, BCs, source_function=lambda x, y, p: 1.0)
    with pytest.raises(TypeError):
        FVM.solve(pois, saveat=0.1)
    with pytest.raises(TypeError):
        FVM.solve(pois, callback=lambda u, t: None)
    assert FVM.solve(pois, reltol=1e-12).resid <= 1e-12


def test_operator_with_columns_beyond_16_bit_offsets(gpu):
    """Nodes numbered in a random order put neighbours far apart: the ELL copy then
    keeps 32-bit columns, and the solvers give the same solution as SuperLU."""
    base = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 200, 190, jitter=0.25, seed=31)
    rng = np.random.default_rng(31)
    perm = rng.permutation(base.num_points())  # new id of each old node
    inv = np.argsort(perm)
    tri = FVM.Triangulation(base.x[inv], base.y[inv], perm[base.triangles], perm[base.boundary_edges],
                            base.boundary_segments)
    m = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(m, lambda x, y, t, u, p: 0.0 * x, FVM.Dirichlet)
    prob = FVM.PoissonsEquation(m, BCs, diffusion_function=lambda x, y, p: 1.0 + 0.5 * np.sin(3 * x) ** 2,
                                source_function=lambda x, y, p: np.cos(2 * x) - y)
    W, n_over, ib = prob.problem.ell_info(index_bytes=True)
    assert ib == 4
    direct = spla.spsolve(prob.A.tocsc(), prob.b)
    for alg in (TM.CG(), TM.MixedCG(), TM.BiCGStab()):
        s = FVM.solve(prob, alg, rtol=1e-11)
        assert s.retcode == "Success", s.stats
        assert np.linalg.norm(s.u - direct) <= 1e-8 * np.linalg.norm(direct)
