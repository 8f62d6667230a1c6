"""Writes the golden fixtures of this directory; run from the repository root as
`python tests/golden/make_golden.py`.

Each fixture is a numpy .npz (no pickled objects) holding one small problem's
mesh, flattened conditions and input state, and the outputs of the CPU oracle
(oracle/, the C restatement of the reference's fvm_eqs! and template assembly):
control volumes and triangle records (geometry.jl:99-169), du of fvm_eqs!
(main_equations.jl:28-81) with its per-node scale sum_sigma |Q_sigma| / V_i + |S_i|,
the Dirichlet update (dirichlet.jl:78-86) and, for the Poisson template, its
assembled CSR operator and right-hand side (abstract_templates.jl:73-325,
poissons_equation.jl:58-88).  Before anything is written, every du is checked
against the brute-force control-volume integration of tests/brute.py
(test_functions.jl:601-646), the control volumes against the domain area, and the
Poisson fixture against its closed form: a fixture records what those
independent checks accept, not just what the oracle printed.

Each fixture names the tests/problems.py builder of the same problem, so a GPU
test rebuilds it through the package and compares the mesh, the flattened
conditions and the state bit for bit before comparing outputs.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "finitevolumemethod.jl_amd"), os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

import finitevolumemethod as FVM  # noqa: E402
from brute import dudt as brute_dudt  # noqa: E402
from oracle import OracleProblem, flatten_conditions, geometry, merge_conditions  # noqa: E402

# example_heat_convection_problem (test_functions.jl:346-374)
HEAT = dict(L=1.0, k=237.0, T0=10.0, Tinf=10.0, alpha=80.0e-6, q=10.0, h=25.0)


def _heat_cond(f, x, y, t, u):
    P = HEAT
    return (-P["alpha"] * P["q"] / P["k"], 0.0, -P["alpha"] * P["h"] / P["k"] * (P["Tinf"] - u), 0.0)[f]


_MIXED_FNS = (lambda x, y, t, u: x + y + t + u + 0.29, lambda x, y, t, u: -u + 3 * t,
              lambda x, y, t, u: x * y + u - 0.5, lambda x, y, t, u: u + 0.2 - t,
              lambda x, y, t, u: 2 * u + 0.25 * t - 0.3, lambda x, y, t, u: y - x)


def _cases():
    """name -> (tests/problems.py builder and its arguments, triangulation, segment
    types, internal conditions, (flux, source, cond), input state, times).  The
    physics restates the builder's closures; the GPU test rebuilds the problem with
    the builder and checks the mesh, the flattened conditions and the state against
    the fixture first."""
    rng = np.random.default_rng
    plate = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, 12, 12, single_boundary=True)
    jit = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, 14, 12, jitter=0.25, seed=3)
    heat = FVM.triangulate_rectangle(0, HEAT["L"], 0, HEAT["L"], 20, 20, single_boundary=False)
    mixed = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 5.0, 12, 19, single_boundary=False)

    def mixed_flux(x, y, t, a, b, g):
        D = 1 + 0.5 * (a * x + b * y + g)
        return -D * a, -D * b

    def nonlinear_flux(x, y, t, a, b, g):
        D = 0.1 + (a * x + b * y + g) ** 2
        return -D * a, -D * b

    zero = lambda f, x, y, t, u: 0.0  # noqa: E731
    return {
        "plate_12": (("diffusion_problem", dict(n=12)), plate, ["Dirichlet"], (0, None, None),
                     (lambda x, y, t, a, b, g: (-a / 9, -b / 9), lambda x, y, t, u: 0.0, zero),
                     np.where(plate.y <= 1.0, 50.0, 0.0), [0.0]),
        "nonlinear_jittered_14x12": (("nonlinear_problem", dict(nx=14, ny=12, seed=3)), jit, ["Dirichlet"],
                                     (0, None, None), (nonlinear_flux, lambda x, y, t, u: u * (1 - u), zero),
                                     rng(4).random(jit.num_points()), [0.0]),
        "heat_convection_20": (("heat_convection_problem", dict(n=20)), heat, ["Neumann"] * 4, (0, None, None),
                               (lambda x, y, t, a, b, g: (-HEAT["alpha"] * a, -HEAT["alpha"] * b),
                                lambda x, y, t, u: 0.0, _heat_cond),
                               HEAT["T0"] + rng(1).random(heat.num_points()), [0.0]),
        "mixed_12x19": (("mixed_problem", dict(nx=12, ny=19, seed=5)), mixed,
                        ["Dirichlet", "Neumann", "Dudt", "Neumann"],
                        (2, {6 + k * 12: 0 for k in range(2, 17)}, {3 + 3 * 12: 1, 4 + 5 * 12: 1}),
                        (mixed_flux, lambda x, y, t, u: u + 1.5, lambda f, x, y, t, u: _MIXED_FNS[f](x, y, t, u)),
                        rng(5).random(mixed.num_points()), [2.0, 3.3]),
    }


def _mesh_arrays(tri):
    return dict(x=tri.x, y=tri.y, triangles=np.asarray(tri.triangles, np.int32),
                boundary_edges=np.asarray(tri.boundary_edges, np.int32))


def rhs_fixture(spec, tri, seg_types, internal, phys, u, times):
    n_int, idir, idudt = internal
    ne, _, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments, seg_types, n_int, idir, idudt)
    kind, fidx, bneu = flatten_conditions(tri.num_points(), tri.boundary_edges, ne, dn, dd)
    badj = tri.boundary_adjacent()
    flux, source, cond = phys
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, badj, kind, fidx, bneu,
                        dict(flux=lambda x, y, t, a, b, g, T: flux(x, y, t, a, b, g),
                             source=lambda x, y, t, uu, n: source(x, y, t, uu), cond=cond))
    cv, props = geometry(tri.x, tri.y, tri.triangles)
    hull = np.array([tri.get_point(int(i)) for i in tri.boundary_edges[:, 0]])
    area = 0.5 * np.sum(hull[:, 0] * np.roll(hull[:, 1], -1) - np.roll(hull[:, 0], -1) * hull[:, 1])
    assert abs(cv.sum() - area) <= 1e-12 * area
    out = dict(_mesh_arrays(tri), node_kind=kind, node_fidx=fidx, bedge_neumann_fidx=bneu, cv=cv, props=props,
               u=u, times=np.array(times, np.float64))
    for k, t in enumerate(times):
        du = orc.rhs(u, t)
        sc = orc.scale(u, t)
        ref, _ = brute_dudt(tri.x, tri.y, tri.triangles, tri.boundary_edges, badj, kind, fidx, bneu,
                            flux, source, cond, u, t)
        bound = 1e-10 * (sc + np.abs(ref)) + 1e-12 * np.max(sc)
        assert np.all(np.abs(du - ref) <= bound), np.max(np.abs(du - ref) / bound)
        out[f"du{k}"], out[f"scale{k}"] = du, sc
        out[f"dirichlet{k}"] = orc.update_dirichlet(u, t)
    out["spec"] = np.array(json.dumps(dict(builder=spec[0], kwargs=spec[1])))
    return out


def poisson_fixture(n=10):
    """poissons_equation.jl:92-99 on an n x n grid: laplace u = -sin(pi x) sin(pi y),
    u = 0 on the boundary of [0, 1]^2."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, n, n, single_boundary=True)
    ne, _, dn, dd = merge_conditions(tri.boundary_edges, tri.boundary_segments, ["Dirichlet"])
    kind, fidx, bneu = flatten_conditions(tri.num_points(), tri.boundary_edges, ne, dn, dd)
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind, fidx, bneu, dict(flux=lambda *a: (0.0, 0.0)))
    f = -np.sin(np.pi * tri.x) * np.sin(np.pi * tri.y)
    rows, cols, vals, b = orc.template_assemble(0, 1.0, f_node=f)
    N = tri.num_points()
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    A.sum_duplicates()
    A.sort_indices()
    u = spla.spsolve(A.tocsc(), b)
    exact = np.sin(np.pi * tri.x) * np.sin(np.pi * tri.y) / (2 * np.pi ** 2)
    err = np.linalg.norm(u - exact) / np.linalg.norm(exact)
    print(f"poisson {n}x{n}: relative L2 error against the closed form {err:.3e}")
    assert err <= 2e-2, err
    bn = tri.boundary_nodes()
    assert np.all(b[bn] == 0.0)
    assert np.array_equal(A[bn].toarray(), np.eye(N)[bn])
    return dict(_mesh_arrays(tri), node_kind=kind, A_indptr=A.indptr.astype(np.int32),
                A_indices=A.indices.astype(np.int32), A_data=A.data, b=b, u=u,
                spec=np.array(json.dumps(dict(builder="poisson", kwargs=dict(n=n)))))


def build_all():
    out = {name: rhs_fixture(*case) for name, case in _cases().items()}
    out["poisson_10"] = poisson_fixture(10)
    return out


def main():
    for name, arrays in build_all().items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrays)
        print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
