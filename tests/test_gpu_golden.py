"""The HIP path against the committed golden fixtures (tests/golden/*.npz), with no
oracle at run time: each problem is rebuilt through the package from the builder
the fixture names, its mesh, flattened conditions and device geometry are compared
with the fixture bit for bit (the s coefficients and CV-edge midpoints; normals,
lengths and volumes to a few ulp), then du of fvm_eqs (the 1e-12 bar, relative to
the stored per-node scale), the Dirichlet update, and the Poisson template's
device-assembled operator (to rounding, within jacobian_sparsity's pattern) and
solution."""
import json

import numpy as np
import pytest
import scipy.sparse as sp

import finitevolumemethod as FVM
import problems as PB
from test_golden import NAMES, load
from test_gpu_rhs import assert_parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", [n for n in NAMES if n != "poisson_10"])
def test_rhs_against_fixture(gpu, name):
    f = load(name)
    spec = json.loads(str(f["spec"]))
    prob = getattr(PB, spec["builder"])(**spec["kwargs"])[0]
    tri = prob.mesh.triangulation
    assert np.array_equal(tri.x, f["x"]) and np.array_equal(tri.y, f["y"])
    assert np.array_equal(np.asarray(tri.triangles, np.int32), f["triangles"])
    assert np.array_equal(np.asarray(tri.boundary_edges, np.int32), f["boundary_edges"])
    for got, key in zip(prob.conditions.flatten(tri), ("node_kind", "node_fidx", "bedge_neumann_fidx")):
        assert np.array_equal(got, f[key]), key
    P = prob.mesh.props_array
    assert np.array_equal(P[:, :15], f["props"][:, :15])
    np.testing.assert_allclose(P[:, 15:], f["props"][:, 15:], rtol=4e-16, atol=4e-16)
    np.testing.assert_allclose(prob.mesh.cv_volumes, f["cv"], rtol=1e-14, atol=0)
    u = f["u"]
    for k, t in enumerate(f["times"]):
        du = FVM.fvm_eqs(np.zeros_like(u), u, prob, float(t))
        assert_parity(du, f[f"du{k}"], f[f"scale{k}"])
        got = FVM.update_dirichlet_nodes(u.copy(), float(t), prob)
        np.testing.assert_allclose(got, f[f"dirichlet{k}"], rtol=1e-15, atol=1e-13)


def test_poisson_template_against_fixture(gpu):
    f = load("poisson_10")
    n = int(json.loads(str(f["spec"]))["kwargs"]["n"])
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, n, n, single_boundary=True)
    assert np.array_equal(tri.x, f["x"]) and np.array_equal(np.asarray(tri.triangles, np.int32), f["triangles"])
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * x, FVM.Dirichlet)
    prob = FVM.PoissonsEquation(mesh, BCs, source_function=lambda x, y, p: -np.sin(np.pi * x) * np.sin(np.pi * y))
    A = prob.A
    A_ref = sp.csr_matrix((f["A_data"], f["A_indices"], f["A_indptr"]), shape=A.shape)
    # the device keeps jacobian_sparsity's pattern: every stored entry lies in it and
    # the values agree to rounding, relative to the largest entry of their row
    rowscale = np.maximum(np.asarray(abs(A_ref).max(axis=1).todense()).ravel(), 1e-300)
    D = (A - A_ref).tocoo()
    assert np.all(np.abs(D.data) <= 1e-13 * rowscale[D.row]), np.max(np.abs(D.data) / rowscale[D.row])
    pattern = sp.csr_matrix((np.ones(A.nnz), A.indices, A.indptr), shape=A.shape)
    ref = A_ref.tocoo()
    assert np.all(np.asarray(pattern[ref.row, ref.col]).ravel() == 1.0)
    np.testing.assert_allclose(prob.b, f["b"], rtol=1e-15, atol=0)
    sol = FVM.solve(prob, rtol=1e-12)
    assert sol.retcode == "Success"
    np.testing.assert_allclose(sol.u, f["u"], rtol=0, atol=1e-11 * np.max(np.abs(f["u"])))
