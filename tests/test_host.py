"""Host-side logic (no GPU): tile invariants, mesh generators, condition merging,
closure tracing, and the C ABI export table."""
import ctypes
import os
import re

import numpy as np
import pytest

import finitevolumemethod as FVM
from finitevolumemethod import _lib, forms
from oracle import merge_conditions

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def tiles(tri, tile_tris=512, max_nodes=640):
    L = _lib.host_lib()
    out = ctypes.POINTER(_lib._Tiles)()
    st = L.fvmh_tiles_build(len(tri.x), _lib.ptr(tri.x), _lib.ptr(tri.y), len(tri.triangles),
                            _lib.ptr(tri.triangles, _lib._i32p), len(tri.boundary_edges),
                            _lib.ptr(tri.boundary_edges, _lib._i32p), tile_tris, max_nodes,
                            ctypes.byref(out))
    assert st == 0, L.fvmh_last_error()
    t = out.contents

    def arr(name, n, dt):
        p = getattr(t, name)
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                     shape=(n,)).copy() if n > 0 else np.zeros(0, dt)

    T, N = t.n_tri, t.n_points
    d = dict(n_tiles=t.n_tiles, n_local=t.n_local, n_fix=t.n_fix, n_parts=t.n_parts)
    d["perm"] = arr("perm", T, np.int32)
    d["tri_sorted"] = arr("tri_sorted", 3 * T, np.int32).reshape(3, T)
    d["tile_tri_start"] = arr("tile_tri_start", t.n_tiles + 1, np.int32)
    d["tile_node_start"] = arr("tile_node_start", t.n_tiles + 1, np.int32)
    d["tile_nodes"] = arr("tile_nodes", t.n_local, np.int32)
    d["loc"] = arr("loc", T, np.uint32)
    d["inc_start"] = arr("inc_start", t.n_local + 1, np.int32)
    d["inc"] = arr("inc", 3 * T, np.uint16)
    d["part_idx"] = arr("part_idx", t.n_local, np.int32)
    d["fix_nodes"] = arr("fix_nodes", t.n_fix, np.int32)
    d["fix_part_start"] = arr("fix_part_start", t.n_fix + 1, np.int32)
    d["fix_bedge_start"] = arr("fix_bedge_start", t.n_fix + 1, np.int32)
    d["fix_bedge"] = arr("fix_bedge", d["fix_bedge_start"][-1], np.int32)
    d["fix_bvert_start"] = arr("fix_bvert_start", t.n_fix + 1, np.int32)
    d["fix_bvert"] = arr("fix_bvert", d["fix_bvert_start"][-1], np.int32)
    d["badj_sorted"] = arr("badj_sorted", t.n_bedges, np.int32)
    L.fvmh_tiles_free(out)
    return d


@pytest.mark.parametrize("tile_tris,max_nodes", [(512, 640), (256, 300), (64, 80), (7, 20)])
def test_tile_invariants(tile_tris, max_nodes):
    base = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 41, 37, jitter=0.25, seed=3)
    # two points that are not vertices
    tri = FVM.Triangulation(np.r_[base.x, 0.5, 0.3], np.r_[base.y, 0.5, 0.2], base.triangles,
                            base.boundary_edges)
    d = tiles(tri, tile_tris, max_nodes)
    N, T = tri.num_points(), len(tri.triangles)
    assert sorted(d["perm"]) == list(range(T))
    assert np.array_equal(d["tri_sorted"].T, tri.triangles[d["perm"]])
    written = np.zeros(N, int)  # exclusive writes + fix-node writes: exactly one per node
    for k in range(d["n_tiles"]):
        t0, t1 = d["tile_tri_start"][k:k + 2]
        n0, n1 = d["tile_node_start"][k:k + 2]
        nt, nn = t1 - t0, n1 - n0
        assert 0 < nt <= tile_tris and 0 < nn <= max_nodes <= 1024
        nodes = d["tile_nodes"][n0:n1]
        assert np.all(np.diff(nodes) > 0)
        loc = d["loc"][t0:t1]
        ids = np.stack([loc & 1023, (loc >> 10) & 1023, (loc >> 20) & 1023], 1)
        assert ids.max() < nn
        assert np.array_equal(nodes[ids], d["tri_sorted"][:, t0:t1].T)
        for ln in range(nn):
            a, b = d["inc_start"][n0 + ln], d["inc_start"][n0 + ln + 1]
            inc = d["inc"][a:b].astype(int)
            assert np.all((inc >> 2) < nt) and np.all((inc & 3) < 3)
            assert np.all(ids[inc >> 2, inc & 3] == ln)
            if d["part_idx"][n0 + ln] < 0:
                written[nodes[ln]] += 1
            else:
                assert d["part_idx"][n0 + ln] < d["n_parts"]
    assert d["inc_start"][-1] == 3 * T
    written[d["fix_nodes"]] += 1
    assert np.all(written == 1)
    fixed = set(d["fix_nodes"].tolist())
    assert {N - 1, N - 2} <= fixed  # points that are not vertices are finished by the fix-up
    assert set(tri.boundary_nodes().tolist()) <= fixed
    assert np.array_equal(np.sort(d["part_idx"][d["part_idx"] >= 0]), np.arange(d["n_parts"]))
    assert d["fix_part_start"][-1] == d["n_parts"]
    be = d["fix_bedge"]
    assert np.all((be >> 1) < len(tri.boundary_edges))
    assert np.all(tri.boundary_edges[be >> 1, be & 1] ==
                  np.repeat(d["fix_nodes"], np.diff(d["fix_bedge_start"])))
    bv = d["fix_bvert"]
    badj = tri.boundary_adjacent()
    assert np.all(tri.triangles[badj[bv >> 2], bv & 3] ==
                  np.repeat(d["fix_nodes"], np.diff(d["fix_bvert_start"])))
    assert np.array_equal(d["perm"][d["badj_sorted"]], badj)


def test_rectangle_numbering_matches_reference_layout():
    # test/equations.jl:42-56: node 201 = (0, h) for nx = 200 and triangle (1, 2, 201)
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 200, 200, single_boundary=False)
    assert tri.get_point(200) == (0.0, 1.0 / 199)
    assert tuple(tri.triangles[0]) == (0, 1, 200)
    assert tri.num_points() == 40000 and len(tri.triangles) == 2 * 199 * 199
    # counter-clockwise triangles
    p = tri.points[tri.triangles]
    d1, d2 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
    area2 = d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0]
    assert np.all(area2 > 0)
    # boundary edges: interior on the left; segments bottom, right, top, left
    e = tri.boundary_edges
    assert np.all(tri.boundary_segments[tri.y[e[:, 0]] + tri.y[e[:, 1]] == 0] == 0)
    assert np.all(tri.boundary_segments[(tri.x[e[:, 0]] == 1) & (tri.x[e[:, 1]] == 1)] == 1)
    assert tri.num_boundary_segments() == 4
    badj = tri.boundary_adjacent()
    assert np.all((tri.triangles[badj] == e[:, :1]).any(1) & (tri.triangles[badj] == e[:, 1:]).any(1))


@pytest.mark.parametrize("n,seed", [(12, 1), (33, 2), (60, 3)])
def test_jittered_mesh_is_the_delaunay_triangulation(n, seed):
    from scipy.spatial import Delaunay

    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, n, n + 3, jitter=0.25, seed=seed)
    assert tri.count_non_delaunay() == 0
    N, B = tri.num_points(), len(tri.boundary_edges)
    assert len(tri.triangles) == 2 * N - 2 - B
    # same triangle set as Qhull (interior points are in general position;
    # boundary cells with collinear boundary points keep a valid Delaunay choice)
    ours = {tuple(sorted(T)) for T in tri.triangles.tolist()}
    qh = Delaunay(tri.points)
    d1 = tri.points[qh.simplices[:, 1]] - tri.points[qh.simplices[:, 0]]
    d2 = tri.points[qh.simplices[:, 2]] - tri.points[qh.simplices[:, 0]]
    area = np.abs(d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0])
    theirs = {tuple(sorted(T)) for T, a in zip(qh.simplices.tolist(), area) if a > 1e-14}
    assert ours == theirs


def test_jitter_bounds_and_seed():
    a = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 30, 30, jitter=0.25, seed=11)
    b = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 30, 30, jitter=0.25, seed=11)
    g = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 30, 30)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.triangles, b.triangles)
    h = 1.0 / 29
    assert np.max(np.abs(a.x - g.x)) <= 0.25 * h and np.max(np.abs(a.y - g.y)) <= 0.25 * h
    bn = g.boundary_nodes()
    assert np.array_equal(a.x[bn], g.x[bn]) and np.array_equal(a.y[bn], g.y[bn])


def test_csr_pattern_matches_brute_force():
    # test_jacobian_sparsity (test_functions.jl:749-763)
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 11, 9, jitter=0.25, seed=1)
    N = tri.num_points()
    A = np.zeros((N, N))
    for i in range(N):
        A[i, i] = 1
        for j in tri.neighbours(i):
            A[i, j] = 1
    J = FVM.jacobian_sparsity(tri)
    assert np.array_equal(J.toarray(), A)


def test_condition_merging_matches_restatement():
    class _Mesh:
        def __init__(self, tri):
            self.triangulation = tri

    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 5.0, 12, 19, single_boundary=False)
    mesh = _Mesh(tri)
    f = lambda x, y, t, u, p: u  # noqa: E731
    types = (FVM.Dirichlet, FVM.Dudt, FVM.Neumann, FVM.Constrained)
    BCs = FVM.BoundaryConditions(mesh, (f,) * 4, types)
    idir = {6 + 12 * k: 0 for k in range(1, 10)}
    idudt = {37 + k: 1 for k in range(8)}
    ICs = FVM.InternalConditions((f, f), dirichlet_nodes=idir, dudt_nodes=idudt)
    conds = FVM.Conditions(mesh, BCs, ICs)
    ne, ce, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments,
                                      [t.name for t in types], 2, idir, idudt)
    assert conds.neumann_edges == ne and conds.constrained_edges == ce
    assert conds.dirichlet_nodes == dn and conds.dudt_nodes == dd
    kind, fidx, bn = conds.flatten(tri)
    for i in range(tri.num_points()):
        if i in dn:
            assert kind[i] == 1 and fidx[i] == dn[i]
        elif i in dd:
            assert kind[i] == 2 and fidx[i] == dd[i]
        else:
            assert kind[i] == 0
    with pytest.raises(AssertionError):
        FVM.BoundaryConditions(mesh, (f,) * 3, types[:3])


def test_diffusion_varying_only_away_from_the_probe_points_is_not_taken_as_constant():
    """The trace runs at fixed probe points inside [0, 1]^2; a D(x, y) that is constant
    there but not over the mesh (the mean-exit-time walkthrough's D, 6.25e-4 inside
    r = R1(theta) ~ 2 and 6.25e-5 outside) is checked at the mesh's points: it becomes the
    evaluated-per-edge form, and a D mixing such (x, y) dependence with u is left to the
    compiled path.  Constant-coefficient closures still pass the check."""
    th = np.linspace(0.0, 2 * np.pi, 200)
    r = np.linspace(1.0, 3.0, 200)
    xs, ys = r * np.cos(th), r * np.sin(th)
    piecewise = lambda x, y, t, u: np.where(np.hypot(x, y) < 2.0, 6.25e-4, 6.25e-5)  # noqa: E731
    assert forms.classify_diffusion(piecewise) == ("poly", [6.25e-4, 0.0, 0.0, 0.0])  # the probes alone
    assert forms.classify_diffusion(piecewise, xs, ys) == ("xy", None)
    with pytest.raises(forms.UnsupportedClosure):
        forms.classify_diffusion(lambda x, y, t, u: np.where(np.hypot(x, y) < 2.0, 1.0, 2.0) * u, xs, ys)
    assert forms.classify_diffusion(lambda x, y, t, u: 0.1 + u ** 2, xs, ys) == ("poly", [0.1, 0.0, 1.0, 0.0])
    assert forms.classify_diffusion(lambda x, y, t, u: 1 / 9 + 0.0 * u, xs, ys) == ("poly", [1 / 9, 0.0, 0.0, 0.0])
    assert forms.classify_diffusion(lambda x, y, t, u: 0.5 * u ** 3 - u + 2.0, xs, ys) == ("poly", [2.0, -1.0, 0.0, 0.5])


def test_closure_tracing_forms():
    assert forms.classify_diffusion(lambda x, y, t, u: 0.1 + u ** 2) == ("poly", [0.1, 0.0, 1.0, 0.0])
    assert forms.classify_diffusion(lambda x, y, t, u: 1 / 9) == ("poly", [1 / 9, 0.0, 0.0, 0.0])
    assert forms.classify_diffusion(lambda x, y, t, u: np.exp(x * y))[0] == "xy"
    with pytest.raises(forms.UnsupportedClosure):
        forms.classify_diffusion(lambda x, y, t, u: np.exp(u))
    with pytest.raises(forms.UnsupportedClosure):
        forms.classify_diffusion(lambda x, y, t, u: 1.0 if u > 0 else 2.0)
    s_node, poly, st = forms.classify_source(lambda x, y, t, u: u * (1 - u), np.zeros(3), np.zeros(3))
    assert s_node is None and poly == [0.0, 1.0, -1.0, 0.0] and st == 0.0
    x = np.array([0.0, 1.0, 2.0])
    r = forms.classify_condition(lambda x, y, t, u: x * y + u - 0.5, x, x + 1)
    np.testing.assert_array_equal(r["g"], x * (x + 1) - 0.5)
    assert (r["cu"], r["ct"], r["c0"], r["lam"]) == (1.0, 0.0, 0.0, 0.0)
    # the annulus tutorial's inner boundary: 50 (1 - exp(-t / 2)) = 50 - 50 exp(-t / 2)
    r = forms.classify_condition(lambda x, y, t, u: 50 * (1 - np.exp(-t / 2)) + 0.0 * u, x, x)
    assert (r["c0"], r["lam"], r["cu"], r["ct"]) == (50.0, -0.5, 0.0, 0.0)
    np.testing.assert_array_equal(r["g"], np.full(3, -50.0))
    with pytest.raises(forms.UnsupportedClosure):
        forms.classify_condition(lambda x, y, t, u: x * t * u, x, x)
    kind, D = forms.classify_flux(lambda x, y, t, a, b, g: (-80e-6 * a, -80e-6 * b))
    assert kind == "xy"
    np.testing.assert_array_equal(D(x, x), np.full(3, 80e-6))


def test_power_law_diffusion_is_recognised():
    """porous_medium_equation.jl:164: D = p.D |u|^(p.m - 1) with m = 3.4 does not
    trace symbolically; it is recognised by evaluation as d0 + c |u|^p.  A
    u^2.4 without abs (complex for u < 0), exp(u) or an (x, y) factor are not."""
    D, m = 0.581, 3.4
    assert forms.classify_diffusion(lambda x, y, t, u: D * abs(u) ** (m - 1)) == ("power", (0.0, D, 2.4))
    assert forms.classify_diffusion(lambda x, y, t, u: 0.1 + 2 * np.abs(u) ** 0.5) == ("power", (0.1, 2.0, 0.5))
    assert forms.classify_diffusion(lambda x, y, t, u: 2.53 * u ** (2 - 1)) == ("poly", [0.0, 2.53, 0.0, 0.0])
    for bad in (lambda x, y, t, u: u ** 2.4, lambda x, y, t, u: np.exp(u),
                lambda x, y, t, u: (1 + x) * np.abs(u) ** 1.5, lambda x, y, t, u: (1 + t) * abs(u) ** 1.5):
        with pytest.raises(forms.UnsupportedClosure):
            forms.classify_diffusion(bad)


def test_linear_flux_forms_trace():
    """piecewise_linear_and_natural_neighbour_interpolation_for_an_advection_diffusion_equation.jl:80-87:
    q = (nu u - D alpha, -D beta) with u = alpha x + beta y + gamma is linear in (alpha,
    beta, gamma) with (x, y)-dependent coefficients; an anisotropic tensor too; a
    t-dependent, a constant or a quadratic term is not."""
    x = np.array([0.1, -0.4, 0.9])
    y = np.array([0.3, 0.8, -0.2])
    D, nu = 0.02, 0.05

    def adv(x, y, t, a, b, g):
        u = a * x + b * y + g
        return nu * u - D * a, -D * b

    kind, A = forms.classify_flux(adv)
    assert kind == "linear"
    got = A(x, y)
    assert got.shape == (6, 3)
    np.testing.assert_array_equal(got[0], nu * x - D)
    np.testing.assert_array_equal(got[1], nu * y)
    np.testing.assert_array_equal(got[2], np.full(3, nu))
    np.testing.assert_array_equal(got[3:5], np.array([[0.0] * 3, [-D] * 3]))
    np.testing.assert_array_equal(got[5], np.zeros(3))
    kind, A = forms.classify_flux(lambda x, y, t, a, b, g: (-(2 + x) * a - 0.5 * b, -0.5 * a - (1 + y ** 2) * b))
    assert kind == "linear"
    np.testing.assert_array_equal(A(x, y)[[0, 1, 3, 4]], np.array([-(2 + x), [-0.5] * 3, [-0.5] * 3, -(1 + y ** 2)]))
    for bad in (lambda x, y, t, a, b, g: (-(1 + t) * a, -b),
                lambda x, y, t, a, b, g: (-a + 1.0, -b),
                lambda x, y, t, a, b, g: (-(a * x + b * y + g) ** 2 * a, -b)):
        with pytest.raises(forms.UnsupportedClosure):
            forms.classify_flux(bad)


def test_c_abi_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "fvm_hip.h")).read()
    declared = set(re.findall(r"^int\s+(fvm_\w+)\s*\(", header, re.M))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(os.path.join(_lib.LIB_DIR, "libfvmhip.so"))
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared <= set(_lib.HIP_API), sorted(declared - set(_lib.HIP_API))


def test_rkc_coefficients_stability_and_order():
    """The RKC(s) recurrence (adjoint.rkc_coefficients) applied to y' = z y: |R(z)| <= 1
    on its stability interval [-beta(s), 0] with beta(s) / s^2 -> 0.653 (damping 2/13),
    second order (local error ~ z^3), R(z) the s-th Chebyshev-based stability
    polynomial; rkc_stages picks the fewest stages covering the requested interval."""
    from finitevolumemethod.adjoint import rkc_coefficients, rkc_stages

    def R(z, c):
        Y0, F0 = 1.0, z
        Ym2, Ym1 = Y0, Y0 + c["mut1"] * F0
        for j in range(2, c["s"] + 1):
            Y = ((1 - c["mu"][j] - c["nu"][j]) * Y0 + c["mu"][j] * Ym1 + c["nu"][j] * Ym2
                 + c["mut"][j] * z * Ym1 + c["gt"][j] * F0)
            Ym2, Ym1 = Ym1, Y
        return Ym1

    for s in (2, 3, 7, 40, 150):
        c = rkc_coefficients(s)
        zs = -np.linspace(0.0, c["beta"], 4001)
        assert max(abs(R(z, c)) for z in zs) <= 1.0 + 1e-12
        # strong damping (eps = 10): stable on its interval, and with many stages the
        # stiff part of it is damped to <= 0.25 per step
        c10 = rkc_coefficients(s, 10.0)
        zs10 = -np.linspace(0.05 * c10["beta"], c10["beta"], 4001)
        r10 = max(abs(R(z, c10)) for z in zs10)
        assert r10 <= 1.0 and (s < 40 or r10 <= 0.25)
        assert abs(R(-1.5 * c["beta"], c)) > 1.0  # the interval ends near beta
        if s >= 40:
            assert abs(c["beta"] / s ** 2 - 0.653) < 2e-3
        e1, e2 = (abs(R(-h, c) - np.exp(-h)) for h in (1e-2, 5e-3))
        assert abs(np.log2(e1 / e2) - 3.0) < 0.05
        assert c["c"][s] == pytest.approx(1.0, abs=1e-12)
    for target in (1.0, 50.0, 5000.0):
        s = rkc_stages(target, margin=1.0)
        assert rkc_coefficients(s)["beta"] >= target
        assert s == 2 or rkc_coefficients(s - 1)["beta"] < target


def test_controller_power_is_the_same_in_python_and_cpp():
    """The step controllers' x^p (fvm_ctrl.hpp, restated in solve.ctrl_pow) built
    from operations that round the same everywhere: the C++ function (compiled for
    the host here; the device build uses the same text) and the Python one agree bit
    for bit, and both are within 1e-14 of x^p."""
    import subprocess
    import tempfile

    from finitevolumemethod.solve import ctrl_pow

    src = ('#include <cmath>\n#include "fvm_ctrl.hpp"\n'
           'extern "C" double cpow(double x, double p) { return fvm::ctrl_pow(x, p); }\n')
    d = tempfile.mkdtemp()
    open(os.path.join(d, "c.cpp"), "w").write(src)
    so = os.path.join(d, "c.so")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC", "-D__host__=",
                    "-D__device__=", "-D__HIPCC_RTC__", "-I", os.path.join(ROOT, "finitevolumemethod.jl_amd", "csrc"),
                    os.path.join(d, "c.cpp"), "-o", so], check=True)
    f = ctypes.CDLL(so).cpow
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(5)
    for x in 10.0 ** rng.uniform(-16, 3, 3000):
        for p in (7 / 50, 2 / 25, 0.8, -1.0 / 3.0):
            a, b = ctrl_pow(float(x), p), f(float(x), p)
            assert a == b, (x, p, a, b)
            assert abs(a - float(x) ** p) <= 1e-14 * float(x) ** p


def test_distance_two_colouring_separates_columns_of_every_row():
    """fvmh_colour_d2 (the colouring behind the Jacobian probes): in every row of
    jacobian_sparsity's pattern the columns carry distinct colours, on a jittered mesh
    and on a triangulate_rectangle grid, with at most 16 colours; a bad column index
    is rejected."""
    from finitevolumemethod.equations import csr_pattern
    from finitevolumemethod.implicit import _colouring

    for tri in (FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 31, 27, jitter=0.25, seed=4),
                FVM.triangulate_rectangle(0.0, 2.0, 0.0, 1.0, 40, 21, single_boundary=True)):
        rp, col = csr_pattern(tri)
        n = tri.num_points()
        c = _colouring(rp, col, n)
        assert c.min() == 0 and c.max() < 16
        for i in range(n):
            cs = c[col[rp[i]:rp[i + 1]]]
            assert len(np.unique(cs)) == len(cs), i
    with pytest.raises(ValueError):
        _colouring(np.array([0, 2, 3]), np.array([0, 5, 1], np.int32), 2)
