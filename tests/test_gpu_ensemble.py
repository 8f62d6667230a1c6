"""Packed ensembles (fvm_batch_*): one launch evaluates every member's fvm_eqs!,
and explicit Euler steps (graph-replayed or not) match stepping each member
alone through the single-problem path."""
import numpy as np
import pytest

import finitevolumemethod as FVM
from finitevolumemethod import ensemble as EN

pytestmark = pytest.mark.gpu


def members(P=6, seed=5):
    """Uneven jittered rectangles on [0, 2]^2; constant D per member except one with
    D(x, y); homogeneous Dirichlet, or Dirichlet + Neumann segments."""
    rng = np.random.default_rng(seed)
    probs = []
    for p in range(P):
        nx, ny = int(rng.integers(8, 41)), int(rng.integers(8, 41))
        single = p % 3 != 2
        tri = FVM.jittered_rectangle(0.0, 2.0, 0.0, 2.0, nx, ny, jitter=0.25, seed=100 + p,
                                     single_boundary=single)
        mesh = FVM.FVMGeometry(tri)
        if single:
            BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
        else:
            BCs = FVM.BoundaryConditions(
                mesh, (lambda x, y, t, u, p: 0.0 * u, lambda x, y, t, u, p: 0.5 + 0.0 * u,
                       lambda x, y, t, u, p: 1.0 + 0.0 * u, lambda x, y, t, u, p: -0.25 * u),
                (FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet, FVM.Neumann))
        D = float(np.exp(np.log(0.1) + 0.5 * rng.standard_normal()))
        if p == 1:
            dfun = lambda x, y, t, u, p: 0.1 * (1 + x * y)  # noqa: E731
        else:
            dfun = (lambda d: (lambda x, y, t, u, p: d + 0.0 * u))(D)
        u0 = rng.random(tri.num_points())
        probs.append(FVM.FVMProblem(mesh, BCs, diffusion_function=dfun, initial_condition=u0,
                                    final_time=1.0))
    return probs


def test_packed_rhs_matches_each_member(gpu):
    probs = members()
    b = EN.DeviceBatch(probs)
    u = b.pack([p.initial_condition for p in probs])
    du = FVM.DeviceArray(b.n)
    b.rhs(FVM.DeviceArray.from_host(u), du, 0.3)
    got = b.unpack(du.to_host())
    for p, g in zip(probs, got):
        ref = FVM.fvm_eqs(np.zeros_like(p.initial_condition), p.initial_condition, p, 0.3)
        scale = np.max(np.abs(ref)) + 1e-300
        assert np.max(np.abs(g - ref)) <= 1e-13 * scale


@pytest.mark.parametrize("graph_steps", [0, 5])
def test_batched_euler_matches_member_loop(gpu, graph_steps):
    probs = members()
    dts = [0.5 * EN.explicit_dt(p) for p in probs]
    nsteps = 23
    b = EN.DeviceBatch(probs, dts, graph_steps=graph_steps)
    d_u = FVM.DeviceArray.from_host(b.pack([p.initial_condition for p in probs]))
    b.euler(d_u, nsteps, 0.0)
    got = b.unpack(d_u.to_host())
    for p, dt, g in zip(probs, dts, got):
        ctx = p.device_context()
        u = FVM.DeviceArray.from_host(p.initial_condition)
        du = FVM.DeviceArray(len(u))
        for k in range(nsteps):
            ctx.rhs(u, du, k * dt)
            FVM._lib.lincomb(u, (1.0, dt), (u, du))
            ctx.apply_dirichlet(u, (k + 1) * dt)
        ref = u.to_host()
        assert np.max(np.abs(g - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_graph_replay_and_direct_steps_agree_bitwise(gpu):
    probs = members(P=4, seed=9)
    dts = [0.5 * EN.explicit_dt(p) for p in probs]
    outs = []
    for gs in (0, 4):
        b = EN.DeviceBatch(probs, dts, graph_steps=gs)
        d_u = FVM.DeviceArray.from_host(b.pack([p.initial_condition for p in probs]))
        b.euler(d_u, 12, 0.0)  # 3 graph replays
        b.euler(d_u, 6, 0.0)   # 1 replay + 2 direct steps
        outs.append(d_u.to_host())
    assert np.array_equal(outs[0], outs[1])


def test_fused_step_equals_rhs_plus_update(gpu):
    """The Euler update fused into the RHS kernels gives the same bits as the RHS
    followed by a separate update kernel."""
    probs = members(P=5, seed=13)
    dts = [0.5 * EN.explicit_dt(p) for p in probs]
    outs = []
    for fused in (True, False):
        b = EN.DeviceBatch(probs, dts, fused=fused)
        d_u = FVM.DeviceArray.from_host(b.pack([p.initial_condition for p in probs]))
        b.euler(d_u, 9, 0.0)
        outs.append(d_u.to_host())
    assert np.array_equal(outs[0], outs[1])


def test_autotuner_reports_candidates(gpu):
    probs = members(P=3, seed=2)
    b = EN.DeviceBatch(probs, autotune=True)
    info = b.info
    assert info.n_candidates == 3
    tiles = list(info.candidate_tiles)[:3]
    assert sorted(tiles) == [256, 512, 1024]
    ms = list(info.candidate_ms)[:3]
    assert all(m > 0 for m in ms)
    assert info.tile_triangles == tiles[int(np.argmin(ms))]


def test_ensemble_solve_and_errors(gpu):
    probs = members(P=3, seed=4)
    dts = [0.5 * EN.explicit_dt(p) for p in probs]
    n = 10
    for p, dt in zip(probs, dts):
        p.final_time = p.initial_time + n * dt
    ens = FVM.EnsembleProblem(probs)
    sol = ens.solve(FVM.Euler(), dt=dts, graph_steps=4, autotune=False)
    assert len(sol) == 3 and sol.stats["steps"] == n
    b = EN.DeviceBatch(probs, dts)
    d_u = FVM.DeviceArray.from_host(b.pack([p.initial_condition for p in probs]))
    b.euler(d_u, n, 0.0)
    for s, ref in zip(sol.u, b.unpack(d_u.to_host())):
        assert np.array_equal(s.u[-1], ref)
    with pytest.raises(ValueError):
        ens.solve(FVM.Euler(), dt=[dts[0], dts[1] / 2, dts[2]])
    # members must share the source polynomial
    q = probs[0]
    other = FVM.FVMProblem(q.mesh, FVM.BoundaryConditions(q.mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet),
                           diffusion_function=lambda x, y, t, u, p: 0.2 + 0.0 * u,
                           source_function=lambda x, y, t, u, p: u, initial_condition=q.initial_condition,
                           final_time=1.0)
    with pytest.raises(FVM.FVMError):
        EN.DeviceBatch([q, other])


def test_ensemble_with_adaptive_algorithm_solves_members_in_turn(gpu):
    """solve(EnsembleProblem(prob; prob_func), Tsit5(); trajectories) solves each member
    on the device with its own adaptive steps; each equals that member solved alone."""
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 15, 15, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    u0 = np.sin(np.pi * tri.x) * np.sin(np.pi * tri.y)
    base = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: 0.1 + 0.0 * u,
                          initial_condition=u0, final_time=0.2)

    def prob_func(p, i, repeat):
        return FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, q, d=0.05 * i: d + 0.0 * u,
                              initial_condition=u0, final_time=0.2)

    ens = FVM.EnsembleProblem(base, prob_func=prob_func)
    sol = FVM.solve(ens, FVM.Tsit5(), trajectories=3, saveat=0.1)
    assert len(sol) == 3 and sol.stats["mode"] == "serial"
    for i, s in enumerate(sol.u):
        alone = FVM.solve(prob_func(base, i + 1, 1), FVM.Tsit5(), saveat=0.1)
        assert len(s.u) == len(alone.u) == 3
        assert all(np.array_equal(a, b) for a, b in zip(s.u, alone.u))
    # decay grows with the member's D
    assert sol.u[0].u[-1].max() > sol.u[1].u[-1].max() > sol.u[2].u[-1].max()


def test_ensemble_of_steady_problems(gpu):
    """The equilibrium-temperature tutorial's pattern (equilibrium_temperature_..._using_ensembleproblems.jl:70-98,
    overview.md:222): an EnsembleProblem over a SteadyFVMProblem whose prob_func varies
    T_inf in the Robin condition k grad T . n = h (T_inf - T), solved member by member;
    each member equals its own steady solve bit for bit, and the temperature rises with
    T_inf."""
    tri = FVM.triangulate_rectangle(0.0, 0.06, 0.0, 0.06, 30, 30, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    k, h = 3.0, 20.0

    def make(Tinf):
        bcs = (lambda x, y, t, T, p: 0.0 * T, lambda x, y, t, T, p: 40.0 + 0.0 * T,
               lambda x, y, t, T, p: -p["h"] * (p["Tinf"] - T) / p["k"], lambda x, y, t, T, p: 70.0 + 0.0 * T)
        BCs = FVM.BoundaryConditions(mesh, bcs, (FVM.Neumann, FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet),
                                     parameters=(None, None, dict(h=h, Tinf=Tinf, k=k), None))
        prob = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, T, p: 1.0 + 0.0 * T,
                              initial_condition=500 * tri.y + 40, final_time=np.inf)
        return FVM.SteadyFVMProblem(prob)

    Tinfs = [10.0, 20.0, 30.0, 40.0]
    ens = FVM.EnsembleProblem(make(Tinfs[0]), prob_func=lambda p, i, repeat: make(Tinfs[i - 1]))
    sol = FVM.solve(ens, trajectories=len(Tinfs))
    assert sol.stats["mode"] == "serial" and len(sol) == len(Tinfs)
    means = []
    for Tinf, s in zip(Tinfs, sol.u):
        assert s.retcode == "Success", s.stats
        single = FVM.solve(make(Tinf))
        assert np.array_equal(s.u[-1], single.u[-1])
        means.append(np.mean(s.u[-1]))
    assert all(b > a for a, b in zip(means, means[1:]))


def test_differing_initial_times_need_time_independent_members(gpu):
    """The packed Euler path starts every member at one time: members that start at
    different times are accepted only when no source or condition depends on t."""
    tri = FVM.jittered_rectangle(0.0, 2.0, 0.0, 2.0, 12, 14, jitter=0.25, seed=31)
    mesh = FVM.FVMGeometry(tri)
    u0 = np.where(tri.y <= 1.0, 50.0, 0.0)

    def make(bc, t0):
        p = FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, bc, FVM.Dirichlet),
                           diffusion_function=lambda x, y, t, u, p: 0.1 + 0.0 * u,
                           initial_condition=u0, initial_time=t0, final_time=t0 + 1.0)
        return p

    steady_bc = lambda x, y, t, u, p: 0.0 * u  # noqa: E731
    moving_bc = lambda x, y, t, u, p: t + 0.0 * u  # noqa: E731
    dt = EN.explicit_dt(make(steady_bc, 0.0))
    n = 8
    for p in (a := [make(steady_bc, 0.0), make(steady_bc, 0.25)]):
        p.final_time = p.initial_time + n * dt
    sol = FVM.EnsembleProblem(a).solve(FVM.Euler(), dt=dt, graph_steps=0, autotune=False)
    assert np.array_equal(sol.u[0].u[-1], sol.u[1].u[-1])  # nothing depends on t
    b = [make(moving_bc, 0.0), make(moving_bc, 0.25)]
    for p in b:
        p.final_time = p.initial_time + n * dt
    with pytest.raises(ValueError):
        FVM.EnsembleProblem(b).solve(FVM.Euler(), dt=dt, graph_steps=0, autotune=False)
