"""Pins of the CPU oracle (no GPU).

What fixes each oracle function, independently of the oracle's own formulas:
  geometry             brute-force control-volume polygon areas; shape functions
                       against numpy.linalg.solve of [x y 1] a = u; centroid /
                       midpoint / normal / length closed forms (test/geometry.jl:10-76)
  fvm_eqs (serial)     brute-force per-node control-volume flux integration
                       (test_functions.jl:601-646) on Dirichlet, all-Neumann, mixed
                       and nonlinear problems; the exact corner-node value of the
                       heat-convection problem (test/equations.jl:113-168);
                       conservation and linear-exactness invariants
  fvm_eqs (threaded)   equals the serial form to rounding (equations.jl:82-91)
  complex step         matches central differences of the real form
  Dirichlet update     a(x_i, y_i, t, u_i) at Dirichlet nodes only
  templates            Poisson closed form (poissons_equation.jl:92-116), mean exit
                       time on a disk (T = (R^2 - r^2) / 4D), the identity
                       A u + b = fvm_eqs!(u) for the linear DiffusionEquation, and
                       the j-row volume quirk of abstract_templates.jl:262
Condition merging (conditions.jl:486-552) is pinned in test_host.py.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import finitevolumemethod as FVM
from brute import control_volumes, dudt as brute_dudt
from oracle import OracleProblem, OracleSystem, flatten_conditions, geometry, merge_conditions
from problems import brusselator_oracle, disk_triangulation, system_conditions


def _disk(nr=12, R=1.0):
    return disk_triangulation(nr, R)


def _meshes():
    return [FVM.triangulate_rectangle(0.0, 2.0, 0.0, 5.0, 12, 19, single_boundary=False),
            FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 17, 14, jitter=0.25, seed=4),
            _disk()]


# ---------------------------------------------------------------------------- geometry
@pytest.mark.parametrize("mesh_idx", [0, 1, 2])
def test_geometry_matches_brute_force_and_closed_forms(mesh_idx):
    tri = _meshes()[mesh_idx]
    cv, P = geometry(tri.x, tri.y, tri.triangles)
    np.testing.assert_allclose(cv, control_volumes(tri.x, tri.y, tri.triangles), rtol=1e-12)
    hull = np.array([tri.get_point(int(i)) for i in tri.boundary_edges[:, 0]])
    area = 0.5 * np.sum(hull[:, 0] * np.roll(hull[:, 1], -1) - np.roll(hull[:, 0], -1) * hull[:, 1])
    assert abs(cv.sum() - area) < 1e-12 * area
    rng = np.random.default_rng(mesh_idx)
    for n, T in enumerate(tri.triangles):
        p = np.stack([tri.x[T], tri.y[T]], 1)
        uu = rng.random(3)
        abc = np.linalg.solve(np.c_[p, np.ones(3)], uu)
        s = P[n, :9]
        got = np.array([s[0:3] @ uu, s[3:6] @ uu, s[6:9] @ uu])
        np.testing.assert_allclose(got, abc, rtol=1e-9, atol=1e-9 * np.abs(abc).max())
        assert abs(s.sum() - 1.0) < 1e-9
        c = p.mean(0)
        for e in range(3):
            m = (p[e] + p[(e + 1) % 3]) / 2
            np.testing.assert_allclose(P[n, 9 + 2 * e: 11 + 2 * e], (c + m) / 2, rtol=1e-14, atol=1e-15)
            ev = c - m
            np.testing.assert_allclose(P[n, 15 + 2 * e: 17 + 2 * e], np.array([ev[1], -ev[0]]) / np.hypot(*ev),
                                       rtol=1e-13, atol=1e-15)
            assert abs(P[n, 21 + e] - np.hypot(*ev)) <= 1e-14 * np.hypot(*ev)


# ---------------------------------------------------------------------------- fvm_eqs
def _problem(tri, seg_types, physics_py, n_internal=0, idir=None, idudt=None, fn_builtin=None):
    ne, ce, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments, seg_types,
                                      n_internal, idir, idudt)
    kind, fidx, bn = flatten_conditions(tri.num_points(), tri.boundary_edges, ne, dn, dd)
    badj = tri.boundary_adjacent()
    flux, source, cond = physics_py
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, badj, kind, fidx, bn,
                        dict(flux=lambda x, y, t, a, b, g, T: flux(x, y, t, a, b, g),
                             source=lambda x, y, t, u, n: source(x, y, t, u), cond=cond))
    return orc, (kind, fidx, bn, badj)


def _check_brute(tri, orc, meta, phys, u, t, rtol=1e-10):
    kind, fidx, bn, badj = meta
    flux, source, cond = phys
    ref, _ = brute_dudt(tri.x, tri.y, tri.triangles, tri.boundary_edges, badj, kind, fidx, bn,
                        flux, source, cond, u, t)
    du = orc.rhs(u, t)
    sc = orc.scale(u, t)
    # brute force solves its own 3x3 systems: where every flux vanishes (constant u)
    # it leaves rounding noise, bounded by a floor relative to the largest scale
    bound = rtol * (sc + np.abs(ref)) + 1e-12 * np.max(sc)
    assert np.all(np.abs(du - ref) <= bound), np.max(np.abs(du - ref) / bound)
    return du


def test_rhs_diffusion_dirichlet_matches_brute_force():
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, 25, 25, single_boundary=True)
    phys = (lambda x, y, t, a, b, g: (-a / 9, -b / 9), lambda x, y, t, u: 0.0,
            lambda f, x, y, t, u: 0.0)
    orc, meta = _problem(tri, ["Dirichlet"], phys)
    u = np.where(tri.y <= 1.0, 50.0, 0.0)
    du = _check_brute(tri, orc, meta, phys, u, 0.0)
    assert np.all(du[tri.boundary_nodes()] == 0.0)


HEAT = dict(k=237.0, T0=10.0, Tinf=10.0, alpha=80.0e-6, q=10.0, h=25.0)


def _heat(n):
    P = HEAT
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, n, n, single_boundary=False)
    walls = [lambda x, y, t, T: -P["alpha"] * P["q"] / P["k"], lambda x, y, t, T: 0.0,
             lambda x, y, t, T: -P["alpha"] * P["h"] / P["k"] * (P["Tinf"] - T),
             lambda x, y, t, T: 0.0]
    phys = (lambda x, y, t, a, b, g: (-P["alpha"] * a, -P["alpha"] * b), lambda x, y, t, u: 0.0,
            lambda f, x, y, t, u: walls[f](x, y, t, u))
    orc, meta = _problem(tri, ["Neumann"] * 4, phys)
    return tri, orc, meta, phys


def test_rhs_all_neumann_matches_brute_force():
    tri, orc, meta, phys = _heat(45)
    u = 10.0 + np.random.default_rng(1).random(tri.num_points())
    _check_brute(tri, orc, meta, phys, u, 0.0)


def test_rhs_exact_corner_node():
    """test/equations.jl:113-168: node 1 of the 200 x 200 heat-convection problem."""
    tri, orc, meta, phys = _heat(200)
    u = np.full(tri.num_points(), HEAT["T0"])
    x1, y1 = tri.x[1], tri.y[200]
    poly = np.array([(0.0, 0.0), (x1 / 2, 0.0), (x1 / 3, y1 / 3), (0.0, y1 / 2)])
    A = 0.5 * abs(np.sum(poly[:, 0] * np.roll(poly[:, 1], -1) - np.roll(poly[:, 0], -1) * poly[:, 1]))
    assert abs(orc.cv[0] - A) < 1e-14 * A
    flux1 = -HEAT["alpha"] * HEAT["q"] / HEAT["k"] * (x1 / 2)  # bottom wall through (0,0)-(x1/2,0)
    flux4 = 0.0 * (y1 / 2)                                     # left wall
    expected = -(1 / A) * (flux1 + flux4)                      # alpha = beta = 0: interior fluxes vanish
    du = orc.rhs(u, 0.0)
    assert abs(du[0] - expected) <= 1e-13 * abs(expected)
    assert abs(orc.rhs(u, 0.0, nthreads=3)[0] - expected) <= 1e-13 * abs(expected)


def test_rhs_mixed_conditions_matches_brute_force():
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 5.0, 12, 19, single_boundary=False)
    fns = [lambda x, y, t, u: x + y + t + u + 0.29, lambda x, y, t, u: -u + 3 * t,
           lambda x, y, t, u: x * y + u - 0.5, lambda x, y, t, u: u + 0.2 - t,
           lambda x, y, t, u: 2 * u + 0.25 * t - 0.3, lambda x, y, t, u: y - x]
    phys = (lambda x, y, t, a, b, g: (-(1 + 0.5 * (a * x + b * y + g)) * a,
                                      -(1 + 0.5 * (a * x + b * y + g)) * b),
            lambda x, y, t, u: u + 1.5, lambda f, x, y, t, u: fns[f](x, y, t, u))
    idir = {6 + 12 * k: 0 for k in range(2, 17)}
    idudt = {3 + 36: 1, 4 + 60: 1}
    orc, meta = _problem(tri, ["Dirichlet", "Neumann", "Dudt", "Neumann"], phys, 2, idir, idudt)
    u = np.random.default_rng(5).random(tri.num_points())
    for t in (2.0, 3.3):
        _check_brute(tri, orc, meta, phys, u, t)


def test_rhs_nonlinear_jittered_matches_brute_force():
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 30, 27, jitter=0.25, seed=9)
    phys = (lambda x, y, t, a, b, g: (-(0.1 + (a * x + b * y + g) ** 2) * a,
                                      -(0.1 + (a * x + b * y + g) ** 2) * b),
            lambda x, y, t, u: u * (1 - u), lambda f, x, y, t, u: 0.0)
    orc, meta = _problem(tri, ["Dirichlet"], phys)
    u = np.random.default_rng(2).random(tri.num_points())
    _check_brute(tri, orc, meta, phys, u, 0.0)


def test_conservation_and_linear_exactness():
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 25, 22, jitter=0.25, seed=3)
    phys = (lambda x, y, t, a, b, g: (-(0.3 + (a * x + b * y + g) ** 2) * a,
                                      -(0.3 + (a * x + b * y + g) ** 2) * b),
            lambda x, y, t, u: 0.0, lambda f, x, y, t, u: 0.0)
    orc, meta = _problem(tri, ["Neumann"], phys)  # zero normal flux everywhere
    u = np.random.default_rng(4).random(tri.num_points())
    du = orc.rhs(u, 0.0)
    assert abs(np.dot(orc.cv, du)) <= 1e-13 * np.dot(orc.cv, orc.scale(u, 0.0))
    phys_lin = (lambda x, y, t, a, b, g: (-0.7 * a, -0.7 * b), lambda x, y, t, u: 0.0,
                lambda f, x, y, t, u: 0.0)
    orc2, meta2 = _problem(tri, ["Dirichlet"], phys_lin)
    ulin = 2 * tri.x - 3 * tri.y + 1
    du2 = orc2.rhs(ulin, 0.0)
    assert np.max(np.abs(du2)) <= 1e-11 * np.max(orc2.scale(ulin, 0.0))


def test_threaded_matches_serial_and_complex_step_matches_differences():
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 40, 35, jitter=0.25, seed=5)
    ne, ce, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments, ["Dirichlet"])
    kind, fidx, bn = flatten_conditions(tri.num_points(), tri.boundary_edges, ne, dn, dd)
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind, fidx, bn, dict(builtin=dict(flux_kind=0, dpoly=[0.1, 0, 1, 0],
                                                          spoly=[0, 1, -1, 0], fn=([0.0], [0.0], [0.0]))))
    u = np.random.default_rng(6).random(tri.num_points())
    ser = orc.rhs(u)
    for nt in (1, 2, 3, 7):
        par = orc.rhs(u, nthreads=nt)
        assert np.all(np.abs(par - ser) <= 1e-13 * orc.scale(u))
    np.testing.assert_array_equal(orc.rhs_cs(u.astype(complex)).real, ser)
    v = np.random.default_rng(7).standard_normal(len(u))
    jv = orc.rhs_cs(u + 1e-30j * v).imag / 1e-30
    h = 1e-6
    fd = (orc.rhs(u + h * v) - orc.rhs(u - h * v)) / (2 * h)
    assert np.max(np.abs(jv - fd)) <= 1e-6 * np.max(np.abs(jv))


def test_dirichlet_update():
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 9, 8, single_boundary=False)
    ne, ce, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments,
                                      ["Dirichlet", "Neumann", "Dirichlet", "Dudt"])
    kind, fidx, bn = flatten_conditions(tri.num_points(), tri.boundary_edges, ne, dn, dd)
    fns = [lambda x, y, t, u: x + 2 * t, None, lambda x, y, t, u: 3 * u - y, lambda x, y, t, u: 0.0]
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind, fidx, bn, dict(flux=lambda *a: (0.0, 0.0),
                                             cond=lambda f, x, y, t, u: fns[f](x, y, t, u)))
    u = np.random.default_rng(0).random(tri.num_points())
    got = orc.update_dirichlet(u, 0.7)
    for i in range(tri.num_points()):
        exp = fns[dn[i]](tri.x[i], tri.y[i], 0.7, u[i]) if i in dn else u[i]
        assert got[i] == pytest.approx(exp, abs=1e-15)


# ---------------------------------------------------------------------------- templates
def _coo(rows, cols, vals, n):
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def test_poisson_template_closed_form():
    """poissons_equation.jl:92-116: laplace(u) = -sin(pi x) sin(pi y), u = 0 on the
    boundary of [0,1]^2 (100 x 100) gives sin(pi x) sin(pi y) / (2 pi^2), rtol 1e-4."""
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 100, 100, single_boundary=True)
    ne, ce, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments, ["Dirichlet"])
    kind, fidx, bn = flatten_conditions(tri.num_points(), tri.boundary_edges, ne, dn, dd)
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind, fidx, bn, dict(flux=lambda *a: (0.0, 0.0)))
    f = -np.sin(np.pi * tri.x) * np.sin(np.pi * tri.y)
    r, c, v, b = orc.template_assemble(0, 1.0, f_node=f)
    u = spla.spsolve(_coo(r, c, v, tri.num_points()).tocsc(), b)
    exact = np.sin(np.pi * tri.x) * np.sin(np.pi * tri.y) / (2 * np.pi ** 2)
    assert np.linalg.norm(u - exact) <= 1e-4 * np.linalg.norm(exact)


def test_mean_exit_time_disk_closed_form():
    """Mean exit time of a disk with D constant: T(r) = (R^2 - r^2) / (4 D)
    (literate_tutorials/mean_exit_time.jl:61-67 with R1 = R2)."""
    tri = _disk(nr=24, R=2.0)
    D = 0.3
    ne, ce, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments, ["Dirichlet"])
    kind, fidx, bn = flatten_conditions(tri.num_points(), tri.boundary_edges, ne, dn, dd)
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind, fidx, bn, dict(flux=lambda *a: (0.0, 0.0)))
    r, c, v, b = orc.template_assemble(1, D)
    T = spla.spsolve(_coo(r, c, v, tri.num_points()).tocsc(), b)
    exact = (4.0 - (tri.x ** 2 + tri.y ** 2)) / (4 * D)
    assert np.linalg.norm(T - exact) <= 5e-3 * np.linalg.norm(exact)


def test_diffusion_template_equals_rhs_operator():
    """A u + b equals fvm_eqs!(u) for the same linear problem: the template's
    triangle / boundary / Dudt terms (abstract_templates.jl:73-160) restate the
    RHS's fluxes with q = -D(x, y) grad u, and a template Neumann value a is the
    RHS Neumann value -D a (literate_wyos/diffusion_equations.jl:405-410)."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 23, 21, jitter=0.2, seed=8, single_boundary=False)
    Dxy = lambda x, y: np.exp(0.5 * np.sin(4 * np.pi * x) * np.cos(4 * np.pi * y))  # noqa: E731
    types = ["Neumann", "Dudt", "Dirichlet", "Neumann"]
    ne, ce, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments, types)
    kind, fidx, bn = flatten_conditions(tri.num_points(), tri.boundary_edges, ne, dn, dd)
    neu = [lambda x, y: 0.3 + x, None, None, lambda x, y: -0.5 * y]
    dudt_fn = lambda x, y: 1.5 - y  # noqa: E731
    cond = lambda f, x, y, t, u: (-Dxy(x, y) * neu[f](x, y) if types[f] == "Neumann"  # noqa: E731
                                  else dudt_fn(x, y) if types[f] == "Dudt" else 0.0)
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind, fidx, bn,
                        dict(flux=lambda x, y, t, a, b, g, T: (-Dxy(x, y) * a, -Dxy(x, y) * b),
                             cond=cond))
    mx, my = _half_points(tri)
    neu_a = np.zeros(2 * len(tri.boundary_edges))
    for e in range(len(tri.boundary_edges)):
        if bn[e] >= 0:
            neu_a[2 * e:2 * e + 2] = neu[bn[e]](mx[2 * e:2 * e + 2], my[2 * e:2 * e + 2])
    r, c, v, b = orc.template_assemble(2, lambda x, y: float(Dxy(x, y)), neu_a=neu_a,
                                       dudt_val=dudt_fn(tri.x, tri.y))
    A = _coo(r, c, v, tri.num_points())
    u = np.random.default_rng(3).random(tri.num_points())
    lhs = A @ u + b
    rhs = orc.rhs(u, 0.0)
    assert np.all(np.abs(lhs - rhs) <= 1e-12 * (orc.scale(u, 0.0) + np.abs(rhs)))


def _half_points(tri):
    be = tri.boundary_edges
    px, py, qx, qy = tri.x[be[:, 0]], tri.y[be[:, 0]], tri.x[be[:, 1]], tri.y[be[:, 1]]
    mijx, mijy = (px + qx) / 2, (py + qy) / 2
    mx = np.empty(2 * len(be))
    my = np.empty(2 * len(be))
    mx[0::2], my[0::2] = (px + mijx) / 2, (py + mijy) / 2
    mx[1::2], my[1::2] = (qx + mijx) / 2, (qy + mijy) / 2
    return mx, my


def test_template_boundary_row_volume_quirk():
    """abstract_templates.jl:262 divides the j row of a non-Neumann boundary edge by
    V_i.  On a Constrained boundary (no node conditions) A u - fvm_eqs!(u) at node
    n is the boundary flux of its incoming edge (prev, n) times (1/V_prev - 1/V_n)."""
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 9, 7, single_boundary=True)
    ne, ce, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments, ["Constrained"])
    kind, fidx, bn = flatten_conditions(tri.num_points(), tri.boundary_edges, ne, dn, dd)
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind, fidx, bn, dict(flux=lambda x, y, t, a, b, g, T: (-2.0 * a, -2.0 * b)))
    r, c, v, b = orc.template_assemble(2, 2.0)
    A = _coo(r, c, v, tri.num_points())
    u = np.random.default_rng(1).random(tri.num_points())
    diff = A @ u + b - orc.rhs(u, 0.0)
    badj = tri.boundary_adjacent()
    mx, my = _half_points(tri)
    for e, (i, j) in enumerate(tri.boundary_edges):
        T = tri.triangles[badj[e]]
        abc = np.linalg.solve(np.c_[tri.x[T], tri.y[T], np.ones(3)], u[T])
        p, q = np.array(tri.get_point(i)), np.array(tri.get_point(j))
        d = q - p
        n = np.array([d[1], -d[0]]) / np.hypot(*d)
        ell = np.hypot(*d) / 2
        flux_j = 2.0 * (abc[0] * n[0] + abc[1] * n[1]) * ell  # D grad u . n at m_j times length
        expected = flux_j * (1 / orc.cv[i] - 1 / orc.cv[j])
        assert diff[j] == pytest.approx(expected, rel=1e-9, abs=1e-12 * abs(flux_j / orc.cv[j]))


# ---------------------------------------------------------------------------- systems
def test_single_variable_system_equals_scalar_problem():
    """An FVMSystem of one variable evaluates exactly as the FVMProblem (the system
    methods of equations/*.jl reduce to the scalar ones)."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 13, 11, jitter=0.2, seed=2, single_boundary=False)
    types = ["Dirichlet", "Neumann", "Dudt", "Neumann"]
    kind, fidx, bn = system_conditions(tri, [types])
    conds = [lambda x, y, t, u: x + u, lambda x, y, t, u: 0.3 * u - t, lambda x, y, t, u: 2.0 * u,
             lambda x, y, t, u: y - 1.0]
    flux = lambda x, y, t, a, b, g: (-(1 + (a * x + b * y + g) ** 2) * a, -(1 + (a * x + b * y + g) ** 2) * b)  # noqa: E731
    src = lambda x, y, t, u: u * (1 - u)  # noqa: E731
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind[0], fidx[0], bn[0],
                        dict(flux=lambda x, y, t, a, b, g, T: flux(x, y, t, a, b, g),
                             source=lambda x, y, t, u, n: src(x, y, t, u),
                             cond=lambda f, x, y, t, u: conds[f](x, y, t, u)))
    sysorc = OracleSystem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                          kind, fidx, bn,
                          lambda x, y, t, a, b, g: tuple(np.array([q]) for q in flux(x, y, t, a[0], b[0], g[0])),
                          lambda v, x, y, t, u: src(x, y, t, u[0]),
                          lambda f, x, y, t, u: conds[f](x, y, t, u[0]))
    u = np.random.default_rng(4).random(tri.num_points())
    assert np.array_equal(sysorc.rhs(u[:, None], 0.4)[:, 0], orc.rhs(u, 0.4))
    assert np.array_equal(sysorc.update_dirichlet(u[:, None], 0.4)[:, 0], orc.update_dirichlet(u, 0.4))


def test_decoupled_system_equals_two_problems():
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 12, 14, jitter=0.2, seed=5, single_boundary=False)
    t1 = ["Dirichlet", "Neumann", "Neumann", "Neumann"]
    t2 = ["Neumann", "Dirichlet", "Dudt", "Neumann"]
    kind, fidx, bn = system_conditions(tri, [t1, t2])
    c1 = [lambda x, y, t, u: 1.0 + x, lambda x, y, t, u: 0.5, lambda x, y, t, u: -0.25 * u, lambda x, y, t, u: y]
    c2 = [lambda x, y, t, u: 0.2 * u, lambda x, y, t, u: 2.0, lambda x, y, t, u: t - u, lambda x, y, t, u: -x]
    allc = c1 + c2
    sysorc = OracleSystem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                          kind, fidx, bn,
                          lambda x, y, t, a, b, g: (np.array([-0.5 * a[0], -(2 + x) * a[1]]),
                                                    np.array([-0.5 * b[0], -(2 + x) * b[1]])),
                          lambda v, x, y, t, u: u[0] - u[0] ** 2 if v == 0 else 0.3 - u[1],
                          lambda f, x, y, t, u: allc[f](x, y, t, u[0] if f < 4 else u[1]))
    u = np.random.default_rng(6).random((tri.num_points(), 2))
    du = sysorc.rhs(u, 0.7)
    for v, (cs, flux, src) in enumerate([(c1, lambda x, a: -0.5 * a, lambda uu: uu - uu ** 2),
                                         (c2, lambda x, a: -(2 + x) * a, lambda uu: 0.3 - uu)]):
        orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                            kind[v], np.where(kind[v] != 0, fidx[v] - 4 * v, -1),
                            np.where(bn[v] >= 0, bn[v] - 4 * v, -1),
                            dict(flux=lambda x, y, t, a, b, g, T, flux=flux: (flux(x, a), flux(x, b)),
                                 source=lambda x, y, t, uu, n, src=src: src(uu),
                                 cond=lambda f, x, y, t, uu, cs=cs: cs[f](x, y, t, uu)))
        assert np.array_equal(du[:, v], orc.rhs(u[:, v], 0.7))


def test_system_dirichlet_update_is_sequential_over_variables():
    """dirichlet.jl:46-53 updates variable 1 at every node, then variable 2 with the
    new values of variable 1 in u[:, i]."""
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 6, 6, single_boundary=True)
    kind, fidx, bn = system_conditions(tri, [["Dirichlet"], ["Dirichlet"]])
    sysorc = OracleSystem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                          kind, fidx, bn, lambda x, y, t, a, b, g: (-a, -b), lambda v, x, y, t, u: 0.0,
                          lambda f, x, y, t, u: 2.0 if f == 0 else u[0] + 1.0)
    u = sysorc.update_dirichlet(np.zeros((tri.num_points(), 2)), 0.0)
    bnodes = tri.boundary_nodes()
    assert np.all(u[bnodes, 0] == 2.0) and np.all(u[bnodes, 1] == 3.0)


def test_brusselator_rhs_matches_exact_time_derivative():
    """At the exact solution, fvm_eqs! equals d/dt of it (-Phi/2, Psi/2) at the free
    nodes, to the discretisation error of a 60 x 60 grid."""
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 60, 60, single_boundary=False)
    orc = brusselator_oracle(tri)
    t = 0.3
    u = np.stack([np.exp(-tri.x - tri.y - t / 2), np.exp(tri.x + tri.y + t / 2)], 1)
    du = orc.rhs(u, t)
    dexact = np.stack([-u[:, 0] / 2, u[:, 1] / 2], 1)
    free = orc.kind.T == 0
    err = np.abs(du - dexact)[free]
    assert np.linalg.norm(err) <= 2e-2 * np.linalg.norm(dexact[free])
    # the Dirichlet rows are zero, and the update puts the exact boundary values in
    assert np.all(du[orc.kind.T == 1] == 0.0)
    z = orc.update_dirichlet(np.zeros_like(u), t)
    d = orc.kind.T == 1
    np.testing.assert_allclose(z[d], u[d], rtol=1e-15)


def test_system_wrapper_reads_transposed_states():
    """States handed over as views of (nv, N) matrices give what C-ordered copies give."""
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 7, 6, single_boundary=False)
    orc = brusselator_oracle(tri)
    U = np.random.default_rng(3).random((2, tri.num_points())) + 0.5
    for t in (0.0, 0.4):
        assert np.array_equal(orc.update_dirichlet(U.T, t), orc.update_dirichlet(U.T.copy(), t))
        assert np.array_equal(orc.rhs(U.T, t), orc.rhs(U.T.copy(), t))
        assert np.array_equal(orc.scale(U.T, t), orc.scale(U.T.copy(), t))


# ---------------------------------------------------------------------------- queries
def test_compute_flux_oracle_is_exact_for_linear_fields():
    """u = a x + b y + c with q = -D grad u: compute_flux = -D (a, b) . n on every edge,
    swapping (i, j) negates it, and boundary edges use their one triangle."""
    from oracle import compute_flux, directed_edge_triangles

    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 2.0, 9, 11, jitter=0.25, seed=4, single_boundary=False)
    cv, props = geometry(tri.x, tri.y, tri.triangles)
    et = directed_edge_triangles(tri.triangles)
    a, b, c, D = 1.5, -0.75, 0.3, 0.7
    u = a * tri.x + b * tri.y + c
    flux = lambda x, y, t, al, be, ga, T: (-D * al, -D * be)  # noqa: E731
    for (i, j) in list(et.keys()):
        ex, ey = tri.x[j] - tri.x[i], tri.y[j] - tri.y[i]
        ell = np.hypot(ex, ey)
        want = -D * (a * ey / ell - b * ex / ell)
        got = compute_flux(tri.x, tri.y, tri.triangles, props, et, u, i, j, 0.0, flux)[0]
        assert got == pytest.approx(want, rel=1e-12, abs=1e-12)
        back = compute_flux(tri.x, tri.y, tri.triangles, props, et, u, j, i, 0.0, flux)[0]
        assert back == pytest.approx(-got, rel=1e-12, abs=1e-12)


def test_compute_flux_oracle_picks_the_triangle_on_the_normal_side():
    """For a nonlinear u the interior-edge flux uses the triangle right of i -> j."""
    from oracle import compute_flux, directed_edge_triangles

    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 5, 5, single_boundary=True)
    cv, props = geometry(tri.x, tri.y, tri.triangles)
    et = directed_edge_triangles(tri.triangles)
    u = tri.x ** 2 + np.sin(3 * tri.y)
    flux = lambda x, y, t, al, be, ga, T: (-al, -be)  # noqa: E731
    (i, j), T = next(((e, T) for e, T in et.items() if (e[1], e[0]) in et))
    Tr = et[(j, i)]
    s = props[Tr]
    va, vb, vc = tri.triangles[Tr]
    al = s[0] * u[va] + s[1] * u[vb] + s[2] * u[vc]
    be = s[3] * u[va] + s[4] * u[vb] + s[5] * u[vc]
    ex, ey = tri.x[j] - tri.x[i], tri.y[j] - tri.y[i]
    ell = np.hypot(ex, ey)
    # the normal (ey, -ex)/ell points into triangle Tr: its centroid lies on the right
    cx, cy = tri.x[tri.triangles[Tr]].mean(), tri.y[tri.triangles[Tr]].mean()
    assert (cx - tri.x[i]) * ey - (cy - tri.y[i]) * ex > 0
    got = compute_flux(tri.x, tri.y, tri.triangles, props, et, u, i, j, 0.0, flux)[0]
    assert got == pytest.approx((ey / ell) * -al + (-ex / ell) * -be, rel=1e-14)


def test_pl_interpolate_oracle_reproduces_linear_fields_and_vertex_values():
    from oracle import directed_edge_triangles, pl_interpolate

    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 8, 7, jitter=0.25, seed=6)
    cv, props = geometry(tri.x, tri.y, tri.triangles)
    et = directed_edge_triangles(tri.triangles)
    rng = np.random.default_rng(2)
    u = np.stack([2.0 * tri.x - tri.y + 0.5, rng.random(tri.num_points())], 1)
    for T in tri.triangles[::3]:
        i, j, k = (int(v) for v in T)
        w = rng.random(3)
        w /= w.sum()
        px, py = w @ tri.x[[i, j, k]], w @ tri.y[[i, j, k]]
        for rot in ((i, j, k), (j, k, i), (k, i, j)):
            v = pl_interpolate(tri.triangles, props, et, u, rot, px, py)
            assert v[0] == pytest.approx(2.0 * px - py + 0.5, rel=1e-13)
            assert v[1] == pytest.approx(w @ u[[i, j, k], 1], rel=1e-12)
        with pytest.raises(KeyError):
            pl_interpolate(tri.triangles, props, et, u, (i, k, j), px, py)
