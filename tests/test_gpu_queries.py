"""GPU parity of compute_flux (fvm_compute_flux, fvm_system_compute_flux) and
pl_interpolate (fvm_pl_interpolate) against the oracle's restatements of
problem.jl:458-487 and utils.jl:23-27, over every directed edge of the mesh (both
orientations of interior edges, boundary edges both ways) and random points of
every triangle in all three rotations.

Bar: within 1e-12 of the oracle relative to |q| (|n_x| + |n_y|) (flux), 1e-13
(interpolant): the device and the oracle run the same operations except D(u) (a
Horner polynomial against the closure) and exp(log D)."""
import numpy as np
import pytest

import finitevolumemethod as FVM
from oracle import compute_flux as orc_flux, directed_edge_triangles, geometry, pl_interpolate as orc_interp
from problems import brusselator_system, dxy_problem, nodal_problem, nonlinear_problem

pytestmark = pytest.mark.gpu


def all_edges(tri):
    et = directed_edge_triangles(tri.triangles)
    e = np.array(sorted(et.keys()), np.int64)
    # boundary edges also in the orientation with the domain on the right
    rev = np.array([(b, a) for (a, b) in et if (b, a) not in et], np.int64)
    return np.concatenate([e, rev]), et


def check_flux(prob, u, flux, nv=None):
    tri = prob.mesh.triangulation
    _, props = geometry(tri.x, tri.y, tri.triangles)
    E, et = all_edges(tri)
    got = FVM.compute_flux(prob, E[:, 0], E[:, 1], u, 0.25)
    U = u.T if nv else u
    ref = np.array([orc_flux(tri.x, tri.y, tri.triangles, props, et, U, int(i), int(j), 0.25, flux)
                    for i, j in E])

    def fabs(x, y, t, a, b, g, T, c):  # (|q|, 0) or (0, |q|): |q| n_x and |q| n_y
        f = np.hypot(*(np.asarray(v, np.float64) for v in flux(x, y, t, a, b, g, T)))
        return (f, 0.0 * f) if c == 0 else (0.0 * f, f)

    mag = sum(np.abs(np.array([orc_flux(tri.x, tri.y, tri.triangles, props, et, U, int(i), int(j), 0.25,
                                        lambda *args, c=c: fabs(*args, c)) for i, j in E]))
              for c in (0, 1))
    ref = ref if nv else ref[:, 0]
    mag = mag if nv else mag[:, 0]
    assert got.shape == ref.shape
    bad = np.abs(got - ref) > 1e-12 * (np.abs(mag) + np.abs(ref)) + 1e-300
    assert not bad.any(), (E[np.nonzero(bad)[0][:5]], np.max(np.abs(got - ref)))
    return E, got


def test_compute_flux_nonlinear_diffusion(gpu):
    prob, orc, u = nonlinear_problem(nx=17, ny=14)

    def flux(x, y, t, a, b, g, T):
        um = a * x + b * y + g
        D = 0.1 + um ** 2
        return -D * a, -D * b

    E, got = check_flux(prob, u, flux)
    # one edge at a time returns a float, as the reference
    v = FVM.compute_flux(prob, int(E[3, 0]), int(E[3, 1]), u, 0.25)
    assert isinstance(v, float) and v == got[3]


def test_compute_flux_diffusion_of_x_y(gpu):
    prob, orc, u = dxy_problem(n=15)
    Dxy = lambda x, y: np.exp(0.5 * np.sin(4 * np.pi * x) * np.cos(4 * np.pi * y))  # noqa: E731
    check_flux(prob, u, lambda x, y, t, a, b, g, T: (-Dxy(x, y) * a, -Dxy(x, y) * b))


def test_compute_flux_nodal_field(gpu):
    prob, orc, u, logd = nodal_problem(nx=13, ny=12)
    tri = prob.mesh.triangulation
    _, props = geometry(tri.x, tri.y, tri.triangles)

    def flux(x, y, t, a, b, g, T):
        s = props[T]
        Dv = np.exp(logd[tri.triangles[T]])
        D = (s[0] * Dv[0] + s[1] * Dv[1] + s[2] * Dv[2]) * x + \
            (s[3] * Dv[0] + s[4] * Dv[1] + s[5] * Dv[2]) * y + (s[6] * Dv[0] + s[7] * Dv[1] + s[8] * Dv[2])
        return -D * a, -D * b

    check_flux(prob, u, flux)


def test_compute_flux_linear_field_is_exact(gpu):
    """u = a x + b y + c, D constant: q . n = -D (a, b) . n on every edge."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 12, 10, jitter=0.25, seed=8)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Neumann)
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 0.7 + 0.0 * u,
                          initial_condition=np.zeros(tri.num_points()), final_time=1.0)
    u = 1.5 * tri.x - 0.75 * tri.y + 0.3
    E, _ = all_edges(tri)
    got = FVM.compute_flux(prob, E[:, 0], E[:, 1], u, 0.0)
    ex, ey = tri.x[E[:, 1]] - tri.x[E[:, 0]], tri.y[E[:, 1]] - tri.y[E[:, 0]]
    ell = np.hypot(ex, ey)
    np.testing.assert_allclose(got, -0.7 * (1.5 * ey / ell + 0.75 * ex / ell), rtol=1e-12, atol=1e-12)


def test_compute_flux_system(gpu):
    """The system branch (a tuple per edge); its first variable equals the scalar
    problem's flux for a decoupled system (test_functions.jl:696-700)."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 11, 9, jitter=0.2, seed=2, single_boundary=False)
    sys = brusselator_system(tri)
    U = np.stack([np.exp(-tri.x - tri.y) + 0.1 * np.sin(5 * tri.y), np.exp(tri.x + tri.y)])
    E, got = check_flux(sys, U, lambda x, y, t, a, b, g, T: (np.array([-a[0] / 4, -a[1] / 4]),
                                                            np.array([-b[0] / 4, -b[1] / 4])), nv=2)
    one = FVM.compute_flux(sys, int(E[0, 0]), int(E[0, 1]), U, 0.25)
    assert isinstance(one, tuple) and len(one) == 2 and one == tuple(got[0])
    mesh = sys.mesh
    BCs = FVM.BoundaryConditions(mesh, (lambda x, y, t, u, p: 0.0 * u,) * 4, (FVM.Neumann,) * 4)
    scalar = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 0.25 + 0.0 * u,
                            initial_condition=U[0], final_time=1.0)
    np.testing.assert_allclose(FVM.compute_flux(scalar, E[:, 0], E[:, 1], U[0], 0.25), got[:, 0],
                               rtol=1e-15, atol=1e-15)


def test_compute_flux_rejects_edges_not_in_the_mesh(gpu):
    prob, orc, u = nonlinear_problem(nx=6, ny=5)
    tri = prob.mesh.triangulation
    far = int(np.argmax(tri.x + tri.y))
    with pytest.raises(FVM.FVMError):
        FVM.compute_flux(prob, 0, far, u, 0.0)
    with pytest.raises(FVM.FVMError):
        FVM.compute_flux(prob, 0, tri.num_points() + 5, u, 0.0)


@pytest.mark.parametrize("nv", [1, 2])
def test_pl_interpolate_matches_oracle(gpu, nv):
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 14, 12, jitter=0.25, seed=5, single_boundary=False)
    _, props = geometry(tri.x, tri.y, tri.triangles)
    et = directed_edge_triangles(tri.triangles)
    rng = np.random.default_rng(nv)
    if nv == 1:
        prob = nonlinear_problem(nx=14, ny=12, seed=5, jitter=0.25)[0]
        tri = prob.mesh.triangulation
        _, props = geometry(tri.x, tri.y, tri.triangles)
        et = directed_edge_triangles(tri.triangles)
        u = rng.random(tri.num_points())
        U = u
    else:
        prob = brusselator_system(tri)
        u = rng.random((2, tri.num_points()))
        U = u.T
    Ts, px, py = [], [], []
    for r, T in enumerate(tri.triangles):
        w = rng.random(3)
        w /= w.sum()
        rot = r % 3
        Ts.append(np.roll(T, -rot))
        px.append(w @ tri.x[T])
        py.append(w @ tri.y[T])
    Ts, px, py = np.array(Ts), np.array(px), np.array(py)
    got = FVM.pl_interpolate(prob, Ts, u, px, py)
    ref = np.array([orc_interp(tri.triangles, props, et, U, T, x, y) for T, x, y in zip(Ts, px, py)])
    ref = ref[:, 0] if nv == 1 else ref
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13)
    one = FVM.pl_interpolate(prob, tuple(Ts[4]), u, px[4], py[4])
    if nv == 1:
        assert isinstance(one, float) and one == got[4]
    else:
        assert np.array_equal(one, got[4])
    with pytest.raises(FVM.FVMError):  # clockwise: not a triangle of the mesh
        FVM.pl_interpolate(prob, (Ts[0][0], Ts[0][2], Ts[0][1]), u, px[0], py[0])


def test_find_triangle_locates_points(gpu):
    """Points drawn strictly inside random triangles of a jittered mesh are found in
    exactly those triangles (also from a fixed start vertex, walking far); points
    outside the rectangle give (-1, -1, -1)."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 40, 35, jitter=0.25, seed=21, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    rng = np.random.default_rng(2)
    T = rng.integers(0, len(tri.triangles), 3000)
    w = rng.uniform(0.05, 1.0, (len(T), 3))
    w /= w.sum(1, keepdims=True)
    V = tri.triangles[T]
    px = np.sum(w * tri.x[V], 1)
    py = np.sum(w * tri.y[V], 1)
    got = FVM.find_triangle(mesh, px, py)
    assert np.array_equal(got, V)
    assert np.array_equal(FVM.find_triangle(mesh, px, py, start=0), V)
    assert FVM.find_triangle(mesh, float(px[7]), float(py[7])) == tuple(int(v) for v in V[7])
    out = FVM.find_triangle(mesh, np.array([-0.1, 0.5, 1.3]), np.array([0.5, 1.2, 0.5]))
    assert np.all(out == -1)


def test_find_triangle_on_a_domain_with_a_hole(gpu):
    """On the annulus 0.2 <= r <= 1 a straight walk from the far side is blocked by the
    hole; such queries are finished by the scan.  Points in the hole are outside."""
    from problems import annulus_triangulation

    tri = annulus_triangulation()
    mesh = FVM.FVMGeometry(tri)
    rng = np.random.default_rng(4)
    T = rng.integers(0, len(tri.triangles), 500)
    w = rng.uniform(0.05, 1.0, (len(T), 3))
    w /= w.sum(1, keepdims=True)
    V = tri.triangles[T]
    px, py = np.sum(w * tri.x[V], 1), np.sum(w * tri.y[V], 1)
    far = int(np.argmin(tri.x))  # a vertex on the left of the hole
    assert np.array_equal(FVM.find_triangle(mesh, px, py, start=far), V)
    th = rng.uniform(0, 2 * np.pi, 50)
    r = rng.uniform(0.0, 0.15, 50)
    assert np.all(FVM.find_triangle(mesh, r * np.cos(th), r * np.sin(th)) == -1)


def test_pl_interpolate_at_located_points(gpu):
    """find_triangle then pl_interpolate, as the advection-diffusion tutorial does with
    jump_and_march: equal to the oracle's interpolant at the same points."""
    prob = nonlinear_problem(nx=20, ny=17, seed=8, jitter=0.25)[0]
    tri = prob.mesh.triangulation
    _, props = geometry(tri.x, tri.y, tri.triangles)
    et = directed_edge_triangles(tri.triangles)
    rng = np.random.default_rng(6)
    px, py = rng.uniform(0.0, 1.0, 400), rng.uniform(0.0, 1.0, 400)
    u = rng.random(tri.num_points())
    Ts = FVM.find_triangle(prob, px, py)
    assert np.all(Ts >= 0)
    got = FVM.pl_interpolate(prob, Ts, u, px, py)
    ref = np.array([orc_interp(tri.triangles, props, et, u, T, x, y)[0] for T, x, y in zip(Ts, px, py)])
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13)
