"""GPU parity of the device geometry and fvm_eqs! against the CPU oracle.

Bars: connectivity and indices bit-exact; geometry fields that need no square
root (s1..s9, CV-edge midpoints) bit-exact, normals and lengths within a few
ulp, control volumes within 1e-14 (different summation order); du within
1e-12 of the oracle in the max norm and per node relative to
sum_sigma |Q_sigma| / V_i + |S_i| (the reference's own threaded and serial forms
differ by rounding of the same size)."""
import numpy as np
import pytest

import finitevolumemethod as FVM
from oracle import geometry as oracle_geometry
from problems import (annulus_problem, diffusion_problem, dxy_problem, heat_convection_problem,
                      mixed_problem, nodal_problem, nonlinear_problem, oracle_for)

pytestmark = pytest.mark.gpu
TOL = 1e-12


def assert_parity(du, ref, scale, tol=TOL):
    du, ref = np.asarray(du), np.asarray(ref)
    assert du.shape == ref.shape
    assert np.all(np.isfinite(du))
    nref = np.max(np.abs(ref))
    assert np.max(np.abs(du - ref)) <= tol * max(nref, 1e-300), np.max(np.abs(du - ref)) / nref
    bound = tol * (scale + 1e-300) + 1e-300
    bad = np.abs(du - ref) > bound
    assert not bad.any(), (np.nonzero(bad)[0][:10], (np.abs(du - ref) / (scale + 1e-300)).max())


@pytest.mark.parametrize("nx,ny,jitter", [(9, 7, 0.0), (40, 33, 0.25), (129, 130, 0.25)])
def test_geometry_matches_oracle(gpu, nx, ny, jitter):
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, nx, ny, jitter=jitter, seed=nx)
    mesh = FVM.FVMGeometry(tri)
    cv_ref, props_ref = oracle_geometry(tri.x, tri.y, tri.triangles)
    P = mesh.props_array
    assert np.array_equal(P[:, :15], props_ref[:, :15])  # s1..s9, CV-edge midpoints
    np.testing.assert_allclose(P[:, 15:], props_ref[:, 15:], rtol=4e-16, atol=4e-16)
    np.testing.assert_allclose(mesh.cv_volumes, cv_ref, rtol=1e-14, atol=0)
    assert abs(mesh.cv_volumes.sum() - 1.0) < 1e-12


def test_tiles_cover_every_node(gpu):
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 70, 61, jitter=0.25, seed=2)
    for tt in (64, 256, 512):
        mesh = FVM.FVMGeometry(tri, tile_triangles=tt)
        assert mesh.info.n_tiles >= len(tri.triangles) // tt


@pytest.mark.parametrize("builder", [nonlinear_problem, diffusion_problem, mixed_problem,
                                     dxy_problem])
def test_fvm_eqs_matches_oracle(gpu, builder):
    prob, orc, u = builder()[:3]
    for t in (prob.initial_time, prob.initial_time + 0.37):
        du = FVM.fvm_eqs(np.zeros_like(u), u, prob, t)
        ref = orc.rhs(u, t)
        assert_parity(du, ref, orc.scale(u, t))


def test_fvm_eqs_heat_convection_neumann(gpu):
    prob, orc, u0 = heat_convection_problem(60)
    u = u0 + np.random.default_rng(0).random(len(u0))
    du = FVM.fvm_eqs(np.zeros_like(u), u, prob, 0.0)
    assert_parity(du, orc.rhs(u, 0.0), orc.scale(u, 0.0))


def test_fvm_eqs_nodal_field(gpu):
    prob, orc, u, logd = nodal_problem()
    du = FVM.fvm_eqs(np.zeros_like(u), u, prob, 0.0)
    assert_parity(du, orc.rhs(u, 0.0), orc.scale(u, 0.0))


def test_fvm_eqs_threaded_oracle_agrees(gpu):
    prob, orc, u = nonlinear_problem(60, 50)
    du = FVM.fvm_eqs(np.zeros_like(u), u, prob, 0.0)
    assert_parity(du, orc.rhs(u, 0.0, nthreads=4), orc.scale(u, 0.0))


def test_points_that_are_not_vertices(gpu):
    base = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 20, 17, jitter=0.25, seed=4)
    x = np.concatenate([base.x, [0.5, 0.25]])
    y = np.concatenate([base.y, [0.5, 0.75]])
    tri = FVM.Triangulation(x, y, base.triangles, base.boundary_edges, base.boundary_segments)
    mesh = FVM.FVMGeometry(tri)
    assert mesh.cv_volumes[-1] == 0 and mesh.cv_volumes[-2] == 0
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    u = np.random.default_rng(1).random(len(x))
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 0.1 + u ** 2,
                          source_function=lambda x, y, t, u, p: u * (1 - u), initial_condition=u,
                          final_time=1.0)
    from problems import oracle_for
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=0, dpoly=[0.1, 0, 1, 0], spoly=[0, 1, -1, 0],
                                             fn=([0.0], [0.0], [0.0]))))
    du = FVM.fvm_eqs(np.zeros_like(u), u, prob, 0.0)
    assert du[-1] == 0 and du[-2] == 0
    assert_parity(du, orc.rhs(u, 0.0), orc.scale(u, 0.0))


@pytest.mark.parametrize("tile", [32, 128, 512])
def test_tile_size_does_not_change_result_beyond_rounding(gpu, tile):
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 50, 45, jitter=0.25, seed=9)
    mesh = FVM.FVMGeometry(tri, tile_triangles=tile)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    u = np.random.default_rng(2).random(tri.num_points())
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 0.1 + u ** 2,
                          source_function=lambda x, y, t, u, p: u * (1 - u), initial_condition=u,
                          final_time=1.0)
    from problems import oracle_for
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=0, dpoly=[0.1, 0, 1, 0], spoly=[0, 1, -1, 0],
                                             fn=([0.0], [0.0], [0.0]))))
    du = FVM.fvm_eqs(np.zeros_like(u), u, prob, 0.0)
    assert_parity(du, orc.rhs(u, 0.0), orc.scale(u, 0.0))


def test_bitwise_deterministic(gpu):
    prob, orc, u = nonlinear_problem(200, 190)
    d_u = FVM.DeviceArray.from_host(u)
    outs = []
    for _ in range(3):
        d_du = FVM.DeviceArray(len(u))
        FVM.fvm_eqs(d_du, d_u, prob, 0.0)
        outs.append(d_du.to_host())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_dirichlet_update_matches_oracle(gpu):
    prob, orc, u = mixed_problem()
    for t in (2.0, 3.5):
        got = FVM.update_dirichlet_nodes(u.copy(), t, prob)
        ref = orc.update_dirichlet(u, t)
        np.testing.assert_allclose(got, ref, rtol=1e-15, atol=1e-15)


def test_invalid_arguments_raise(gpu):
    prob, orc, u = nonlinear_problem(10, 9)
    d = FVM.DeviceArray.from_host(u)
    with pytest.raises(FVM.FVMError):
        prob.device_context().rhs(d, d, 0.0)  # du aliasing u is rejected by the C ABI


@pytest.mark.parametrize("x0,y0", [(0.0, 0.0), (1000.0, -500.0)])
def test_rebuilt_record_matches_stored_geometry(gpu, x0, y0):
    """The record the tile kernels rebuild from vertex coordinates (quotients by
    Delta and 3 from one reciprocal plus Markstein's correction) equals the stored
    TriangleProperties bit for bit in s1..s9 and the CV-edge midpoints, also far from
    the origin where Delta cancels; the unnormalised normals equal n l."""
    import re

    from finitevolumemethod import _lib

    tri = FVM.jittered_rectangle(x0, x0 + 1.0, y0, y0 + 1.0, 257, 233, jitter=0.25, seed=17)
    mesh = FVM.FVMGeometry(tri)
    P = mesh.props_array
    R = np.empty_like(P)
    _lib.check(_lib.hip_lib().fvm_mesh_rebuilt_geometry(mesh.handle, _lib.ptr(R)), "fvm_mesh_rebuilt_geometry")
    assert np.array_equal(R[:, :15], P[:, :15])
    level = int(re.search(r"geo_level=(\d)", _lib.build_info()).group(1))
    if level >= 4:
        nl = P[:, 15:21] * np.repeat(P[:, 21:24], 2, axis=1)
        np.testing.assert_allclose(R[:, 15:21], nl, rtol=1e-15, atol=1e-15 * np.abs(nl).max())
        assert np.all(R[:, 21:24] == 1.0)
    else:
        assert np.array_equal(R[:, 15:], P[:, 15:])


@pytest.mark.parametrize("variant", ["tutorial", "exp"])
def test_conditions_with_exponentials_in_time(gpu, variant):
    """Conditions c0 + cu u + ct t + g(x, y) exp(lam t) on a domain with a hole: the
    annulus tutorial's inner Dirichlet 50 (1 - exp(-t/2))
    (diffusion_equation_on_an_annulus.jl:55), and Neumann / Dudt conditions of that
    form, against the oracle calling the closures themselves."""
    prob, orc, u0 = annulus_problem(variant)
    u = u0 + np.random.default_rng(3).random(len(u0))
    for t in (0.0, 0.7, 1.9):
        du = FVM.fvm_eqs(np.zeros_like(u), u, prob, t)
        assert_parity(du, orc.rhs(u, t), orc.scale(u, t))
        got = FVM.update_dirichlet_nodes(u.copy(), t, prob)
        np.testing.assert_allclose(got, orc.update_dirichlet(u, t), rtol=1e-15, atol=1e-13)


def test_annulus_trajectory_matches_oracle(gpu):
    """SSPRK3 (dt = 2e-4, 150 steps) with the Dirichlet callback after every step:
    the inner boundary follows 50 (1 - exp(-t/2)) and every node the oracle's
    trajectory stepped by the same scheme."""
    dt, nsteps = 2e-4, 150
    prob, orc, u0 = annulus_problem("tutorial", final_time=dt * nsteps)
    sol = FVM.solve(prob, FVM.SSPRK3(dt=dt))
    tri = prob.mesh.triangulation
    inner = np.unique(tri.boundary_edges[tri.boundary_segments == 1])
    u, t = u0.copy(), 0.0
    for _ in range(nsteps):
        u1 = u + dt * orc.rhs(u, t)
        u2 = 0.75 * u + 0.25 * u1 + 0.25 * dt * orc.rhs(u1, t + dt)
        u = 1 / 3 * u + 2 / 3 * u2 + 2 / 3 * dt * orc.rhs(u2, t + 0.5 * dt)
        t = t + dt
        u = orc.update_dirichlet(u, t)
    assert len(sol.u) == nsteps + 1 and sol.t[-1] == pytest.approx(t, abs=1e-15)
    np.testing.assert_allclose(sol.u[-1][inner], 50 * (1 - np.exp(-t / 2)), rtol=1e-13)
    assert np.max(np.abs(sol.u[-1] - u)) <= 1e-11 * np.max(np.abs(u))


def test_points_outside_the_triangulation(gpu):
    """Points that belong to no triangle (source_contributions.jl:36, has_vertex; the
    templates' fix_missing_vertices!, abstract_templates.jl:317-325): their du is 0,
    the other nodes' du matches the oracle at the 1e-12 bar and equals the du of the
    same mesh without the extra points to the same bar; the Poisson template gives
    them identity rows (solution 0) and leaves the solution elsewhere unchanged."""
    base = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 23, 19, jitter=0.25, seed=12)
    n0 = base.num_points()
    extra_x, extra_y = np.array([0.3, 2.0, -1.0]), np.array([0.7, 2.0, 0.5])  # inside and outside the domain
    ext = FVM.Triangulation(np.r_[base.x, extra_x], np.r_[base.y, extra_y], base.triangles, base.boundary_edges,
                            base.boundary_segments)
    rng = np.random.default_rng(3)
    u_ext = rng.random(ext.num_points())
    dus = {}
    for name, tri, u in (("ext", ext, u_ext), ("base", base, u_ext[:n0])):
        mesh = FVM.FVMGeometry(tri)
        BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
        prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 0.1 + u ** 2,
                              source_function=lambda x, y, t, u, p: u * (1 - u), initial_condition=u,
                              final_time=1.0)
        du = FVM.fvm_eqs(np.zeros_like(u), u, prob, 0.0)
        if name == "ext":
            orc = oracle_for(prob, dict(builtin=dict(flux_kind=0, dpoly=[0.1, 0, 1, 0], spoly=[0, 1, -1, 0],
                                                     fn=([0.0], [0.0], [0.0]))))
            assert_parity(du, orc.rhs(u, 0.0), orc.scale(u, 0.0))
            assert np.all(du[n0:] == 0.0)
            assert np.all(mesh.cv_volumes[n0:] == 0.0)
            scale = orc.scale(u, 0.0)[:n0]
        dus[name] = du
    assert_parity(dus["ext"][:n0], dus["base"], scale)
    sols = {}
    for name, tri in (("ext", ext), ("base", base)):
        mesh = FVM.FVMGeometry(tri)
        BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * x, FVM.Dirichlet)
        prob = FVM.PoissonsEquation(mesh, BCs, source_function=lambda x, y, p: np.cos(3 * x) + y)
        sols[name] = FVM.solve(prob, rtol=1e-12)
        assert sols[name].retcode == "Success"
    assert np.all(sols["ext"].u[n0:] == 0.0)
    np.testing.assert_allclose(sols["ext"].u[:n0], sols["base"].u, rtol=0, atol=1e-10 * np.abs(sols["base"].u).max())
