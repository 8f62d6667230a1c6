"""The benchmark workloads at their full sizes, checked against the oracle.

* Ensemble: the 256-member workload (nx, ny ~ logU[20, 300] on [0, 2]^2, D_p ~
  LogNormal(log 0.1, 0.5), u0 = 50 where y <= 1), 20 explicit Euler steps replayed
  from a captured HIP graph; every member against the oracle's fvm_eqs! stepped by
  the same dt followed by update_dirichlet_nodes! (equations/dirichlet.jl:78-86),
  and the packed steps against each member stepped alone, bit for bit.
* Differentiation: the 1000 x 1000 per-node log-D workload; the VJP in u by the
  adjoint identity against the JVP, the JVP against the oracle's complex step, the
  log-D VJP against the oracle's complex-step J_logD z, and both against central
  differences with step 1e-6 (bar 1e-6).
* Steady templates: the 1000 x 1000 PoissonsEquation and MeanExitTimeProblem
  workload; every device solver's true residual |b - A x| / |b| <= 1e-10 on the
  oracle's own assembled A and b (specific_problems/abstract_templates.jl:73-325).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import finitevolumemethod as FVM
from finitevolumemethod import ensemble as EN
from finitevolumemethod import templates as TM
from finitevolumemethod import workloads as WL
from problems import oracle_for, template_oracle

pytestmark = pytest.mark.gpu


def _dev(a):
    return FVM.DeviceArray.from_host(np.asarray(a, np.float64))


@pytest.fixture(scope="module")
def ensemble():
    probs, Ds = WL.ensemble_members(256)
    dts = [EN.explicit_dt(p) for p in probs]
    return probs, Ds, dts


def test_ensemble_workload_shape(gpu, ensemble):
    probs, Ds, dts = ensemble
    sizes = WL.ensemble_sizes(256)
    assert len(probs) == 256
    assert all(20 <= nx <= 300 and 20 <= ny <= 300 for nx, ny, _ in sizes)
    for p, (nx, ny, D), d in zip(probs, sizes, Ds):
        tri = p.mesh.triangulation
        assert tri.num_points() == nx * ny and D == d
        np.testing.assert_array_equal(p.initial_condition, np.where(tri.y <= 1.0, 50.0, 0.0))
    assert all(dt > 0 for dt in dts)


def test_ensemble_graph_steps_match_oracle_and_member_loop(gpu, ensemble):
    probs, Ds, dts = ensemble
    nsteps, tile = 20, 512
    b = EN.DeviceBatch(probs, dts, tile_triangles=tile, graph_steps=10)
    d_u = _dev(b.pack([p.initial_condition for p in probs]))
    b.euler(d_u, nsteps, 0.0)  # two replays of a 10-step graph
    packed = b.unpack(d_u.to_host())
    worst = 0.0
    for k, (p, D, dt, g) in enumerate(zip(probs, Ds, dts, packed)):
        orc = oracle_for(p, dict(builtin=dict(flux_kind=0, dpoly=[D, 0, 0, 0], fn=([0.0], [0.0], [0.0]))))
        u = np.array(p.initial_condition, np.float64)
        for s in range(nsteps):
            u = orc.update_dirichlet(u + dt * orc.rhs(u, s * dt), (s + 1) * dt)
        err = np.max(np.abs(g - u)) / np.max(np.abs(u))
        worst = max(worst, err)
        assert err <= 1e-12, (k, err)
    print(f"ensemble: worst member error {worst:.2e} of max|u| after {nsteps} steps")
    # each member stepped alone through the single-problem path, same tile shape
    for k in range(0, 256, 17):
        one = EN.DeviceBatch([probs[k]], [dts[k]], tile_triangles=tile, graph_steps=0)
        d1 = _dev(probs[k].initial_condition)
        one.euler(d1, nsteps, 0.0)
        assert np.array_equal(d1.to_host(), packed[k]), k


@pytest.fixture(scope="module")
def differentiation():
    tri, prob, u, logd, w = WL.differentiation_problem(1000)
    orc = oracle_for(prob, dict(builtin=dict(flux_kind=1, logD=logd, fn=([0.0], [0.0], [0.0]))))
    return tri, prob, u, logd, w, orc


def test_differentiation_workload_vjp_u(gpu, differentiation):
    tri, prob, u, logd, w, orc = differentiation
    assert tri.num_points() == 1_000_000 and len(tri.triangles) == 1_996_002
    ctx = prob.device_context()
    d_u, d_w = _dev(u), _dev(w)
    ub, lb = FVM.DeviceArray(len(u)), FVM.DeviceArray(len(u))
    ctx.vjp(d_u, d_w, ub, lb, 0.0)
    ub = ub.to_host()
    rng = np.random.default_rng(21)
    jv = FVM.DeviceArray(len(u))
    du_p, du_m = FVM.DeviceArray(len(u)), FVM.DeviceArray(len(u))
    h = 1e-6
    for k in range(3):
        v = rng.standard_normal(len(u))
        ctx.jvp(d_u, _dev(v), jv, 0.0)
        Jv = jv.to_host()
        if k == 0:  # the tangent itself against the oracle's complex step
            ref = orc.rhs_cs(u + 1j * 1e-30 * v, logd, 0.0).imag / 1e-30
            assert np.max(np.abs(Jv - ref)) <= 1e-11 * np.max(np.abs(ref))
        lhs, rhs = np.dot(w, Jv), np.dot(ub, v)
        assert abs(lhs - rhs) <= 1e-11 * np.sum(np.abs(w * Jv))
        ctx.rhs(_dev(u + h * v), du_p, 0.0)
        ctx.rhs(_dev(u - h * v), du_m, 0.0)
        fd = np.dot(w, (du_p.to_host() - du_m.to_host()) / (2 * h))
        assert abs(fd - rhs) <= 1e-6 * max(1.0, abs(fd)), (fd, rhs)


def test_differentiation_workload_vjp_logd(gpu, differentiation):
    tri, prob, u, logd, w, orc = differentiation
    ctx = prob.device_context()
    d_u, d_w = _dev(u), _dev(w)
    ub, lb = FVM.DeviceArray(len(u)), FVM.DeviceArray(len(u))
    ctx.vjp(d_u, d_w, ub, lb, 0.0)
    lb = lb.to_host()
    rng = np.random.default_rng(22)
    du_p, du_m = FVM.DeviceArray(len(u)), FVM.DeviceArray(len(u))
    h = 1e-6
    try:
        for k in range(3):
            z = rng.standard_normal(len(u))
            got = np.dot(lb, z)
            if k == 0:  # ForwardDiff's role: complex step through the oracle in log D
                ref = orc.rhs_cs(u.astype(complex), logd + 1j * 1e-30 * z, 0.0).imag / 1e-30
                assert abs(got - np.dot(w, ref)) <= 1e-11 * np.sum(np.abs(w * ref))
            ctx.set_log_d(_dev(logd + h * z))
            ctx.rhs(d_u, du_p, 0.0)
            ctx.set_log_d(_dev(logd - h * z))
            ctx.rhs(d_u, du_m, 0.0)
            fd = np.dot(w, (du_p.to_host() - du_m.to_host()) / (2 * h))
            assert abs(fd - got) <= 1e-6 * max(1.0, abs(fd)), (fd, got)
    finally:
        ctx.set_log_d(_dev(logd))


@pytest.fixture(scope="module")
def templates():
    tri, pois, met, f = WL.template_problems(1000)
    kind, _, _ = pois.conditions.flatten(tri)
    out = {}
    for name, prob, k, kw in (("poisson", pois, 0, dict(f_node=np.where(kind == 1, 0.0, f))),
                              ("mean_exit_time", met, 1, {})):
        r, c, v, b = template_oracle(prob).template_assemble(k, "exp_sincos:1.0", **kw)
        n = prob.problem.n
        out[name] = (prob, sp.csr_matrix((v, (r, c)), shape=(n, n)), b)
    return tri, out


@pytest.mark.parametrize("alg", [TM.CG(), TM.MixedCG(), TM.MixedCG(update="restart"), TM.BiCGStab()],
                         ids=["cg", "mixed", "mixed_restart", "bicgstab"])
@pytest.mark.parametrize("which", ["poisson", "mean_exit_time"])
def test_template_workload_true_residual(gpu, templates, which, alg):
    tri, out = templates
    assert tri.num_points() == 1_000_000
    prob, A, b = out[which]
    sol = FVM.solve(prob, alg, rtol=1e-10)
    assert sol.retcode == "Success", sol.stats
    rel = np.linalg.norm(b - A @ sol.u) / np.linalg.norm(b)
    assert rel <= 1e-10 * 1.05, (rel, sol.stats)
