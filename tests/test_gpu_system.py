"""GPU parity of FVMSystem (fvm_system_rhs, fvm_system_dirichlet_apply) against the
oracle's serial system restatement, plus the Brusselator's exact solution.

Bars: du within 1e-12 of the oracle in the max norm and per entry relative to
sum_sigma |Q_sigma| / V_i + |S_i| (as for the scalar path); the Dirichlet update
within 1e-14 relative (the device evaluates g exp(lam t) where the closure
evaluates exp of the sum); a two-variable system that decouples into copies of one
FVMProblem equal to that problem's RHS bit for bit."""
import numpy as np
import pytest

import finitevolumemethod as FVM
from oracle import OracleSystem
from problems import (brusselator_oracle, brusselator_system, gray_scott_oracle, gray_scott_system,
                      system_conditions)

pytestmark = pytest.mark.gpu
TOL = 1e-12


def assert_parity(du, ref, scale, tol=TOL):
    du, ref, scale = np.asarray(du), np.asarray(ref), np.asarray(scale)
    assert du.shape == ref.shape
    assert np.all(np.isfinite(du))
    nref = np.max(np.abs(ref))
    assert np.max(np.abs(du - ref)) <= tol * max(nref, 1e-300), np.max(np.abs(du - ref)) / nref
    bad = np.abs(du - ref) > tol * scale + 1e-300
    assert not bad.any(), (np.argwhere(bad)[:10], (np.abs(du - ref) / (scale + 1e-300)).max())


def exact(tri, t):
    return np.stack([np.exp(-tri.x - tri.y - t / 2), np.exp(tri.x + tri.y + t / 2)])


@pytest.mark.parametrize("tile", [256, 512, 1024])
def test_brusselator_rhs_matches_oracle(gpu, tile):
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 47, 41, jitter=0.2, seed=7, single_boundary=False)
    sys = brusselator_system(tri, tile_triangles=tile)
    orc = brusselator_oracle(tri)
    rng = np.random.default_rng(tile)
    for t in (0.0, 0.3, 1.7):
        U = exact(tri, t) * (1 + 0.2 * rng.standard_normal((2, tri.num_points())))
        du = FVM.fvm_eqs(np.zeros_like(U), U, sys, t)
        assert_parity(du, orc.rhs(U.T, t).T, orc.scale(U.T, t).T)


def test_brusselator_dirichlet_update_matches_oracle(gpu):
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 23, 29, jitter=0.2, seed=3, single_boundary=False)
    sys = brusselator_system(tri)
    orc = brusselator_oracle(tri)
    U = np.random.default_rng(1).random((2, tri.num_points()))
    for t in (0.0, 0.8):
        got = FVM.update_dirichlet_nodes(U.copy(), t, sys)
        ref = orc.update_dirichlet(U.T, t).T
        d = orc.kind == 1
        np.testing.assert_allclose(got[d], ref[d], rtol=1e-14, atol=0)
        assert np.array_equal(got[~d], U[~d])


def test_decoupled_system_equals_scalar_problem(gpu):
    """example_diffusion_problem_system (test_functions.jl:318-334): two copies of
    example_diffusion_problem as a system; each row of du is the scalar du."""
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, 25, 25, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u[0] * 0.0 * u[1], FVM.Dirichlet)
    u0 = np.where(tri.y <= 1.0, 50.0, 0.0)
    p1 = FVM.FVMProblem(mesh, BCs, flux_function=lambda x, y, t, a, b, g, p: (-a[0] / 9, -b[0] / 9),
                        initial_condition=u0, final_time=0.5)
    p2 = FVM.FVMProblem(mesh, BCs, flux_function=lambda x, y, t, a, b, g, p: (-a[1] / 9, -b[1] / 9),
                        initial_condition=u0, final_time=0.5)
    sys = FVM.FVMSystem(p1, p2)
    scalar = FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet),
                            diffusion_function=lambda x, y, t, u, p: 1 / 9 + 0.0 * u,
                            initial_condition=u0, final_time=0.5)
    rng = np.random.default_rng(2)
    U = np.stack([u0, 40 * rng.random(len(u0))])
    du = FVM.fvm_eqs(np.zeros_like(U), U, sys, 0.0)
    for v in range(2):
        assert np.array_equal(du[v], FVM.fvm_eqs(np.zeros(len(u0)), U[v], scalar, 0.0))
    assert sys.initial_condition.shape == (2, tri.num_points())


def _three_variable(tile_triangles=0, final_time=1.0):
    """Three variables with cross-diffusion, u-dependent and exponential-in-time
    conditions on every kind of boundary, internal Dudt / Dirichlet nodes, and
    sources with s(x, y), monomials of every variable and a t term."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 31, 27, jitter=0.2, seed=9, single_boundary=False)
    mesh = FVM.FVMGeometry(tri, tile_triangles=tile_triangles)
    N = tri.num_points()
    interior = np.setdiff1d(np.arange(N), tri.boundary_nodes())
    na, nb = int(interior[len(interior) // 3]), int(interior[2 * len(interior) // 3])
    c0 = [lambda x, y, t, u: x + 2 * t, lambda x, y, t, u: 0.5 * u[0] - u[2] + y,
          lambda x, y, t, u: u[1] - 3 * t, lambda x, y, t, u: np.exp(-t) * y]
    c1 = [lambda x, y, t, u: 0.0 * u[0], lambda x, y, t, u: 1 + u[0],
          lambda x, y, t, u: -0.2 * u[1], lambda x, y, t, u: 2.0 + np.exp(x + 0.5 * t)]
    c2 = [lambda x, y, t, u: 0.1 * x, lambda x, y, t, u: 0.3 * u[2] - t,
          lambda x, y, t, u: 0.0 * u[0], lambda x, y, t, u: 0.25 + 0.0 * t]
    i1 = lambda x, y, t, u: 0.5 * u[1] - u[0]  # noqa: E731  Dudt at node na for variable 1
    i2 = lambda x, y, t, u: 3.0 + x * y + t  # noqa: E731  Dirichlet at node nb for variable 2
    K = np.array([[1.0, 0.2, 0.0], [0.0, 0.5, 0.1], [0.3, 0.0, 2.0]])
    flux = [lambda x, y, t, a, b, g, p, v=v: (-(K[v, 0] * a[0] + K[v, 1] * a[1] + K[v, 2] * a[2]),
                                              -(K[v, 0] * b[0] + K[v, 1] * b[1] + K[v, 2] * b[2]))
            for v in range(3)]
    src = [lambda x, y, t, u: 1 + x * y - u[0] * u[1] + 0.5 * t,
           lambda x, y, t, u: u[0] * u[1] * u[2] - u[1] ** 3,
           lambda x, y, t, u: 2 * u[2] - 0.1]
    types = [(FVM.Dirichlet, FVM.Neumann, FVM.Dudt, FVM.Neumann),
             (FVM.Neumann, FVM.Dirichlet, FVM.Neumann, FVM.Dudt),
             (FVM.Neumann, FVM.Neumann, FVM.Neumann, FVM.Neumann)]
    internal = [FVM.InternalConditions(),
                FVM.InternalConditions((lambda x, y, t, u, p: i1(x, y, t, u),), dudt_nodes={na: 0}),
                FVM.InternalConditions((lambda x, y, t, u, p: i2(x, y, t, u),), dirichlet_nodes={nb: 0})]
    rng = np.random.default_rng(5)
    probs = []
    for v, cs in enumerate((c0, c1, c2)):
        BCs = FVM.BoundaryConditions(mesh, tuple(lambda x, y, t, u, p, c=c: c(x, y, t, u) for c in cs),
                                     types[v])
        probs.append(FVM.FVMProblem(mesh, BCs, internal[v], flux_function=flux[v],
                                    source_function=lambda x, y, t, u, p, s=src[v]: s(x, y, t, u),
                                    initial_condition=rng.random(N), final_time=final_time))
    sys = FVM.FVMSystem(*probs)
    # oracle: boundary functions 4 v + s, then 12 (i1) and 13 (i2)
    kind, fidx, bn = system_conditions(tri, [[ty.name for ty in tv] for tv in types])
    kind[1, na], fidx[1, na] = 2, 12
    kind[2, nb], fidx[2, nb] = 1, 13
    allc = c0 + c1 + c2 + [i1, i2]
    orc = OracleSystem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                       kind, fidx, bn,
                       lambda x, y, t, a, b, g: (-(K @ a), -(K @ b)),
                       lambda v, x, y, t, u: src[v](x, y, t, u),
                       lambda f, x, y, t, u: allc[f](x, y, t, u))
    return tri, sys, orc


@pytest.mark.parametrize("tile", [0, 256])
def test_three_variable_system_matches_oracle(gpu, tile):
    tri, sys, orc = _three_variable(tile)
    rng = np.random.default_rng(11)
    for t in (0.0, 0.45):
        U = rng.random((3, tri.num_points())) + 0.5
        du = FVM.fvm_eqs(np.zeros_like(U), U, sys, t)
        assert_parity(du, orc.rhs(U.T, t).T, orc.scale(U.T, t).T)
        got = FVM.update_dirichlet_nodes(U.copy(), t, sys)
        ref = orc.update_dirichlet(U.T, t).T
        np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15)


def test_three_variable_euler_steps_match_oracle(gpu):
    """solve(sys, Euler(dt)) against the oracle's loop u <- C(u + dt fvm_eqs!(u), t + dt)."""
    dt, n = 2e-5, 25  # inside the explicit stability limit (D up to 2, h ~ 1/30)
    tri, sys, orc = _three_variable(final_time=n * dt)
    sol = FVM.solve(sys, FVM.Euler(dt=dt))
    assert len(sol.u) == n + 1
    U = sys.initial_condition.T.copy()
    for k in range(n):
        U = orc.update_dirichlet(U + dt * orc.rhs(U, k * dt), (k + 1) * dt)
    assert sol.u[-1].shape == (3, tri.num_points())
    np.testing.assert_allclose(sol.u[-1], U.T, rtol=1e-11, atol=1e-12 * np.abs(U).max())


def test_brusselator_solution_tracks_exact_solution(gpu):
    """The tutorial's check: the solved system against Phi = exp(-x - y - t/2),
    Psi = exp(x + y + t/2) (reaction_diffusion_brusselator_system_of_pdes.jl)."""
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 40, 40, single_boundary=False)
    sys = brusselator_system(tri, final_time=1.0)
    sol = FVM.solve(sys, FVM.Tsit5(), saveat=0.5, abstol=1e-9, reltol=1e-9)
    assert sol.retcode == "Success"
    assert [round(t, 12) for t in sol.t] == [0.0, 0.5, 1.0]
    for t, U in zip(sol.t, sol.u):
        ex = exact(tri, t)
        for v in range(2):
            err = np.linalg.norm(U[v] - ex[v]) / np.linalg.norm(ex[v])
            print(f"t={t} var={v} relative L2 error {err:.3e}")
            assert err <= 5e-4, (t, v)  # measured: 1.3e-4 at t = 1


def test_gray_scott_rhs_matches_oracle(gpu):
    """The device sums the source's monomials (0.04 - 0.04 u + ...) where the closure
    evaluates b (1 - u) - u v^2, so the scale adds the monomials' magnitudes: the
    rounding bound of either evaluation."""
    sys = gray_scott_system(n=60)
    tri = sys.mesh.triangulation
    orc = gray_scott_oracle(tri)
    U = sys.initial_condition
    du = FVM.fvm_eqs(np.zeros_like(U), U, sys, 0.0)
    u, v = U
    mono = np.stack([0.04 + 0.04 * np.abs(u) + np.abs(u * v * v), 0.1 * np.abs(v) + np.abs(u * v * v)])
    assert_parity(du, orc.rhs(U.T, 0.0).T, orc.scale(U.T, 0.0).T + mono)


def test_system_jacobian_sparsity_holds_the_jacobian(gpu):
    """Every nonzero of the finite-difference Jacobian of the system RHS lies in
    jacobian_sparsity(sys) (solve.jl:96-131)."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 7, 6, jitter=0.2, seed=1, single_boundary=False)
    sys = brusselator_system(tri)
    S = FVM.jacobian_sparsity(sys).toarray() != 0
    U = exact(tri, 0.2)
    f0 = FVM.fvm_eqs(np.zeros_like(U), U, sys, 0.2).T.ravel()
    n = U.size
    for c in range(n):
        Up = U.T.ravel().copy()
        Up[c] += 1e-6
        fc = FVM.fvm_eqs(np.zeros_like(U), Up.reshape(-1, 2).T, sys, 0.2).T.ravel()
        nz = np.nonzero(fc != f0)[0]
        assert S[nz, c].all(), (c, nz[~S[nz, c]])


def test_steady_system_solves_to_a_zero_residual(gpu):
    """SteadyFVMProblem(FVMSystem): Newton with GMRES on fvm_system_jvp; at the answer
    the oracle's du vanishes at the free nodes."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 21, 19, jitter=0.2, seed=4)
    mesh = FVM.FVMGeometry(tri)
    B0 = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u[0], FVM.Dirichlet)
    B1 = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: x + 0.0 * u[0], FVM.Dirichlet)
    N = tri.num_points()
    p0 = FVM.FVMProblem(mesh, B0, flux_function=lambda x, y, t, a, b, g, p: (-a[0], -b[0]),
                        source_function=lambda x, y, t, u, p: 1.0 - u[0] + 0.3 * u[1],
                        initial_condition=np.zeros(N), final_time=np.inf)
    p1 = FVM.FVMProblem(mesh, B1, flux_function=lambda x, y, t, a, b, g, p: (-0.5 * a[1], -0.5 * b[1]),
                        source_function=lambda x, y, t, u, p: 0.5 - u[1] + 0.2 * u[0],
                        initial_condition=np.zeros(N), final_time=np.inf)
    sys = FVM.FVMSystem(p0, p1)
    sol = FVM.solve(FVM.SteadyFVMProblem(sys))
    assert sol.retcode == "Success"
    U = sol.u[-1]
    assert U.shape == (2, N)
    kind, fidx, bn = system_conditions(tri, [["Dirichlet"], ["Dirichlet"]])
    orc = OracleSystem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                       kind, fidx, bn,
                       lambda x, y, t, a, b, g: (np.array([-a[0], -0.5 * a[1]]), np.array([-b[0], -0.5 * b[1]])),
                       lambda v, x, y, t, u: 1.0 - u[0] + 0.3 * u[1] if v == 0 else 0.5 - u[1] + 0.2 * u[0],
                       lambda f, x, y, t, u: 0.0 if f == 0 else x)
    r = orc.rhs(U.T, 0.0)
    sc = orc.scale(U.T, 0.0)
    assert np.linalg.norm(r) <= 1e-7 * np.linalg.norm(sc)
    bnd = tri.boundary_nodes()
    assert np.all(U[0, bnd] == 0.0) and np.allclose(U[1, bnd], tri.x[bnd], rtol=0, atol=0)


def _system_jvp(sys, U, Vd, t):
    ctx = sys.device_context()
    d_u = FVM.DeviceArray.from_host(sys.flat(U))
    d_v = FVM.DeviceArray.from_host(sys.flat(Vd))
    d_jv = FVM.DeviceArray(U.size)
    ctx.jvp(d_u, d_v, d_jv, t)
    return sys.unflat(d_jv.to_host())


def test_system_jvp_matches_central_differences(gpu):
    """fvm_system_jvp against central differences of the oracle's system RHS (step
    1e-5: truncation ~1e-10 of the scale) on the Brusselator and the three-variable
    system (sources of degree 3, u-dependent Dudt / Neumann / Dirichlet conditions)."""
    rng = np.random.default_rng(21)
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 19, 17, jitter=0.2, seed=7, single_boundary=False)
    cases = [(brusselator_system(tri), brusselator_oracle(tri), 0.3)]
    t3, s3, o3 = _three_variable()
    cases.append((s3, o3, 0.45))
    for sys, orc, t in cases:
        nv, N = sys.initial_condition.shape
        U = rng.random((nv, N)) + 0.5
        Vd = rng.standard_normal((nv, N))
        jv = _system_jvp(sys, U, Vd, t)
        h = 1e-5
        fd = (orc.rhs((U + h * Vd).T, t) - orc.rhs((U - h * Vd).T, t)).T / (2 * h)
        scale = orc.scale(U.T, t).T * (1 + np.abs(Vd)) + np.abs(fd)
        assert np.all(np.abs(jv - fd) <= 1e-6 * scale + 1e-12), np.max(np.abs(jv - fd) / scale)
        # Dirichlet entries and their derivative are zero
        assert np.all(jv[orc.kind == 1] == 0.0)


def test_system_jvp_is_exact_for_a_linear_system(gpu):
    """F(u) = A u + b (linear fluxes, affine sources and conditions): J v = F(v) - F(0)
    to rounding."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 15, 13, jitter=0.2, seed=3, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    N = tri.num_points()
    B0 = FVM.BoundaryConditions(mesh, (lambda x, y, t, u, p: 0.5 * u[1] + x, lambda x, y, t, u, p: 1.0 + 0.0 * u[0],
                                       lambda x, y, t, u, p: -0.25 * u[0], lambda x, y, t, u, p: 0.0 * u[0]),
                                (FVM.Neumann, FVM.Dirichlet, FVM.Dudt, FVM.Neumann))
    B1 = FVM.BoundaryConditions(mesh, (lambda x, y, t, u, p: 0.0 * u[0], lambda x, y, t, u, p: u[0] - u[1],
                                       lambda x, y, t, u, p: y + 0.0 * u[0], lambda x, y, t, u, p: 2.0 * u[1]),
                                (FVM.Dirichlet, FVM.Neumann, FVM.Neumann, FVM.Dudt))
    p0 = FVM.FVMProblem(mesh, B0, flux_function=lambda x, y, t, a, b, g, p: (-(a[0] + 0.3 * a[1]), -(b[0] + 0.3 * b[1])),
                        source_function=lambda x, y, t, u, p: 1.0 - u[0] + 0.3 * u[1],
                        initial_condition=np.zeros(N), final_time=1.0)
    p1 = FVM.FVMProblem(mesh, B1, flux_function=lambda x, y, t, a, b, g, p: (-0.5 * a[1], -0.5 * b[1]),
                        source_function=lambda x, y, t, u, p: 0.5 - u[1] + 0.2 * u[0] + x * y,
                        initial_condition=np.zeros(N), final_time=1.0)
    sys = FVM.FVMSystem(p0, p1)
    rng = np.random.default_rng(4)
    U = rng.random((2, N))
    Vd = rng.standard_normal((2, N))
    jv = _system_jvp(sys, U, Vd, 0.2)
    F = lambda W: FVM.fvm_eqs(np.zeros_like(W), W, sys, 0.2)  # noqa: E731
    ref = F(Vd) - F(np.zeros_like(Vd))
    mag = np.abs(F(Vd)) + np.abs(F(np.zeros_like(Vd))) + 1.0
    assert np.all(np.abs(jv - ref) <= 1e-12 * mag * 1e3)


def test_system_c_abi_rejects_bad_arguments(gpu):
    """fvm_system_* and the query entry points turn bad input into FVM_ERR_ARG with a
    message (no kernel runs on it): aliasing, unknown variables counts, out-of-range
    function indices and source exponents, and edges / triangles not in the mesh."""
    import ctypes

    from finitevolumemethod import _lib
    from finitevolumemethod.system import system_desc, system_physics

    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 9, 8, jitter=0.2, seed=1, single_boundary=False)
    sys = brusselator_system(tri)
    ctx = sys.device_context()
    d = FVM.DeviceArray(2 * tri.num_points())
    with pytest.raises(FVM.FVMError):
        ctx.rhs(d, d, 0.0)
    with pytest.raises(FVM.FVMError):
        ctx.jvp(d, d, d, 0.0)
    L = _lib.hip_lib()
    ph = system_physics(sys)

    def create(**changes):
        bad = dict(ph)
        bad.update(changes)
        desc, keep = system_desc(bad)
        h = ctypes.c_void_p()
        st = L.fvm_system_create(sys.mesh.handle, ctypes.byref(desc), ctypes.byref(h))
        if st == 0:
            L.fvm_system_destroy(h)
        return st, _lib.last_error()

    assert create()[0] == 0
    st, msg = create(n_vars=5, flux_k=np.eye(5))
    assert st != 0 and "n_vars" in msg
    fidx = ph["node_fidx"].copy()
    fidx[ph["node_kind"] != 0] = 99
    st, msg = create(node_fidx=fidx)
    assert st != 0 and "node_fidx" in msg
    st, msg = create(terms=[(0, (4, 0), 1.0)])
    assert st != 0 and "exponents" in msg
    with pytest.raises(FVM.FVMError, match="not in the mesh"):
        FVM.compute_flux(sys, 0, tri.num_points() - 1, sys.initial_condition, 0.0)
    with pytest.raises(FVM.FVMError, match="not in the mesh"):
        FVM.pl_interpolate(sys, (0, 1, 2 * tri.num_points()), sys.initial_condition, 0.1, 0.1)
