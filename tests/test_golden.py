"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) on the CPU.  The generator is deterministic and
checks every fixture against pins independent of the oracle before writing it
(brute-force control-volume integration, the domain area, the Poisson closed
form); here the fixtures must be exactly what it writes today, so a change to
the oracle, the mesh generators or the condition merging shows up as a fixture
that no longer matches, and the stored data are pinned again without the oracle."""
import glob
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz")))


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixture_set():
    assert NAMES == ["heat_convection_20", "mixed_12x19", "nonlinear_jittered_14x12", "plate_12", "poisson_10"]


@pytest.fixture(scope="module")
def regenerated():
    import sys
    sys.path.insert(0, HERE)
    import make_golden
    return make_golden.build_all()


@pytest.mark.parametrize("name", NAMES)
def test_generator_reproduces_fixture(regenerated, name):
    stored, fresh = load(name), regenerated[name]
    assert sorted(stored) == sorted(fresh)
    for k, v in stored.items():
        assert np.array_equal(v, np.asarray(fresh[k])), (name, k)


@pytest.mark.parametrize("name", [n for n in NAMES if n != "poisson_10"])
def test_stored_rhs_data_pins(name):
    """Without the oracle: the control volumes tile the domain, Dirichlet nodes have
    du = 0 and keep every non-Dirichlet value under the update, and nodes without a
    condition and away from the boundary see du = 0 when u is constant."""
    f = load(name)
    x, y, be = f["x"], f["y"], f["boundary_edges"]
    hull = np.stack([x[be[:, 0]], y[be[:, 0]]], 1)
    area = 0.5 * np.sum(hull[:, 0] * np.roll(hull[:, 1], -1) - np.roll(hull[:, 0], -1) * hull[:, 1])
    assert abs(f["cv"].sum() - area) <= 1e-12 * area
    assert np.all(f["cv"] > 0)
    dirichlet = f["node_kind"] == 1
    for k in range(len(f["times"])):
        assert np.all(f[f"du{k}"][dirichlet] == 0.0)
        assert np.array_equal(f[f"dirichlet{k}"][~dirichlet], f["u"][~dirichlet])
        assert np.all(np.isfinite(f[f"du{k}"])) and np.all(f[f"scale{k}"] >= 0)
    spec = json.loads(str(f["spec"]))
    assert spec["builder"] in ("diffusion_problem", "nonlinear_problem", "heat_convection_problem", "mixed_problem")


def test_stored_poisson_pins():
    """The stored operator solves to the stored u, which is second-order close to
    sin(pi x) sin(pi y) / (2 pi^2) (1.0e-2 at 10 x 10), with identity rows at the
    Dirichlet boundary and weakly diagonally dominant free rows."""
    f = load("poisson_10")
    n = len(f["x"])
    A = sp.csr_matrix((f["A_data"], f["A_indices"], f["A_indptr"]), shape=(n, n))
    u = spla.spsolve(A.tocsc(), f["b"])
    np.testing.assert_allclose(u, f["u"], rtol=1e-12, atol=1e-15)
    exact = np.sin(np.pi * f["x"]) * np.sin(np.pi * f["y"]) / (2 * np.pi ** 2)
    assert np.linalg.norm(u - exact) <= 1.1e-2 * np.linalg.norm(exact)
    fixed = f["node_kind"] == 1
    assert np.array_equal(A[np.nonzero(fixed)[0]].toarray(), np.eye(n)[fixed])
    free = np.nonzero(~fixed)[0]
    d = A.diagonal()[free]
    off = np.asarray(abs(A[free]).sum(axis=1)).ravel() - np.abs(d)
    assert np.all(np.abs(d) >= off * (1 - 1e-12))
