"""Benchmark of the FP64 fvm_eqs! right-hand side on one MI355X.

Workload (the reference's throughput configuration): a 2000 x 2000-point grid on
[0, 1]^2 (4,000,000 nodes), interior points jittered by U(-h/4, h/4) per axis
(SplitMix64, seed 20261202) and Delaunay re-triangulated (7,992,002 triangles),
homogeneous Dirichlet boundary, D(u) = 0.1 + u^2, S(u) = u (1 - u), u ~ U[0, 1)
i.i.d.  One step = one fvm_eqs!(du, u, p, 0) over the whole mesh with u and du
resident in HBM.  Throughput = node updates per second.

Printed JSON (one line): the metric; the roofline of the two RHS kernels on the
step time, bound by VALU issue (VALU instructions per RHS and HBM bytes per RHS from
the committed rocprofv3 counter runs of this build, profiles/fvm_rhs_4M_counters.json),
with the HBM fraction and the reference data model's canonical bytes beside it; a
copy-kernel bandwidth; the VJP/RHS cost ratio on the 1000 x 1000 nodal-log-D mesh and
the log-D fit; and the CPU baseline: the oracle's OpenMP restatement of
parallel_fvm_eqs! timed on a bounded number of evaluations.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "finitevolumemethod.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "FP64 fvm_eqs! Gnode-updates/s @4M nodes (%HBM peak); VJP/RHS ratio"
HBM_PEAK = 8.0e12  # B/s, MI355X
# FP64 FMA peak of the vector units, MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2 flops x
# 2.4 GHz.  Every wave64 VALU instruction (FP64, FP32 or integer) holds its SIMD for the
# 4 cycles of an FP64 FMA, so this is also the VALU issue ceiling when each VALU
# instruction is counted as an FMA's 64 x 2 flops.
VALU_PEAK = 78.6e12
SEED = 20261202
COUNTERS = os.path.join(ROOT, "profiles", "fvm_rhs_4M_counters.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def canonical_bytes(n_nodes, n_tri, n_bedges, per_tri=204, per_node=25, per_bedge=44):
    """Algorithmic bytes of one RHS: every array of the reference's data model read
    once and du written once (triangle record 204 B, node 25 B, boundary edge 44 B)."""
    return per_tri * n_tri + per_node * n_nodes + per_bedge * n_bedges


def build_throughput_problem(n_side, tile=0):
    from finitevolumemethod import workloads as WL

    t0 = time.time()
    tri, mesh, prob, u = WL.throughput_problem(n_side, tile)
    log(f"mesh {tri.num_points()} nodes {len(tri.triangles)} triangles, FVMGeometry on device: "
        f"{mesh.info.n_tiles} tiles, {mesh.info.n_fix_nodes} fix nodes ({time.time() - t0:.1f}s)")
    prob.device_context()
    return tri, mesh, prob, u


def rhs_timed(ctx, d_u, d_du):
    from finitevolumemethod import _lib

    a, b = ctypes.c_float(), ctypes.c_float()
    _lib.check(_lib.hip_lib().fvm_rhs_timed(ctx.handle, d_u.ptr, d_du.ptr, 0.0, None,
                                            ctypes.byref(a), ctypes.byref(b)), "fvm_rhs_timed")
    return a.value, b.value


def copy_bandwidth(n=1 << 27, reps=20):
    """The library's vector kernel as a copy (out = 1.0 * x, 16-byte loads and stores,
    two pairs of doubles per thread and iteration) of n doubles: bytes moved / time, the
    achievable HBM rate on this box."""
    from finitevolumemethod import _lib

    x = _lib.DeviceArray(n)
    y = _lib.DeviceArray(n)
    x.fill_zero()
    _lib.lincomb(y, (1.0,), (x,))
    tm = _lib.Timer()
    tm.start()
    for _ in range(reps):
        _lib.lincomb(y, (1.0,), (x,))
    ms = tm.stop()
    return 2 * 8 * n * reps / (ms * 1e-3)


def vjp_ratio(n_side, reps=20, cold_reps=10):
    """Time fvm_rhs and fvm_vjp (u and log D) on the nodal-log-D workload: warm
    (back-to-back calls, the ~0.45 GB working set partly resident in the 256 MB MALL)
    and cold (each call timed alone after writing a 1 GiB buffer, which evicts the MALL
    and the L2s)."""
    from finitevolumemethod import _lib
    from finitevolumemethod import workloads as WL

    tri, prob, u, logd, w = WL.differentiation_problem(n_side)
    ctx = prob.device_context()
    d_u, d_w = _lib.DeviceArray.from_host(u), _lib.DeviceArray.from_host(w)
    d_du, d_ub, d_lb = (_lib.DeviceArray(len(u)) for _ in range(3))
    for _ in range(3):
        ctx.rhs(d_u, d_du)
        ctx.vjp(d_u, d_w, d_ub, d_lb)
    tm = _lib.Timer()
    tm.start()
    for _ in range(reps):
        ctx.rhs(d_u, d_du)
    t_rhs = tm.stop() / reps
    tm.start()
    for _ in range(reps):
        ctx.vjp(d_u, d_w, d_ub, d_lb)
    t_vjp = tm.stop() / reps
    flush = _lib.DeviceArray(1 << 27)

    def cold(fn):
        ts = []
        for _ in range(cold_reps):
            flush.fill_zero()
            tm.start()
            fn()
            ts.append(tm.stop())
        return float(np.mean(ts))

    c_rhs = cold(lambda: ctx.rhs(d_u, d_du))
    c_vjp = cold(lambda: ctx.vjp(d_u, d_w, d_ub, d_lb))
    del flush
    return dict(nodes=int(tri.num_points()), triangles=int(len(tri.triangles)),
                rhs_ms=round(t_rhs, 4), vjp_ms=round(t_vjp, 4), ratio=round(t_vjp / t_rhs, 3),
                cold_rhs_ms=round(c_rhs, 4), cold_vjp_ms=round(c_vjp, 4), cold_ratio=round(c_vjp / c_rhs, 3))


def trbdf2_bench(n_side=1000, tf=0.05):
    """TR-BDF2 (the host-driven integrator: Newton with GMRES(30) on fvm_jvp, each
    Arnoldi step one library call) on the throughput workload's problem at n_side^2
    nodes (D = 0.1 + u^2, S = u (1 - u), homogeneous Dirichlet, u0 ~ U[0, 1)) from
    t = 0 to tf, saving at tf, with GMRES right-preconditioned by diag(W)^-1 (the
    default) and without: GMRES iterations per Newton step for both."""
    import finitevolumemethod as FVM
    from finitevolumemethod import workloads as WL

    tri, mesh, prob, u = WL.throughput_problem(n_side)
    prob.final_time = tf
    out = dict(workload=f"TRBDF2 on the throughput problem at {n_side}x{n_side} from t = 0 to {tf:g}",
               nodes=int(tri.num_points()))
    for pc in ("jacobi", "none"):
        t0 = time.perf_counter()
        sol = FVM.solve(prob, FVM.TRBDF2(precond=pc), saveat=[tf])
        sec = time.perf_counter() - t0
        st = sol.stats
        out[f"precond_{pc}"] = dict(
            retcode=sol.retcode, seconds=round(sec, 3),
            accepted_steps=int(st.get("naccept", 0)), rejected_steps=int(st.get("nreject", 0)),
            newton_iterations=int(st.get("nnewton", 0)), gmres_iterations=int(st.get("nkrylov", 0)),
            newton_failures=int(st.get("nnewton_fail", 0)),
            gmres_per_newton=round(st.get("nkrylov", 0) / max(1, st.get("nnewton", 1)), 1))
    log(f"TR-BDF2 {out}")
    return out


def fit_bench(n_side, iters=50, tf=0.1, nsteps=10):
    """The diffusivity-field fit of the differentiation workload (n_side x n_side jittered
    mesh, u0 ~ U[0, 1), homogeneous Dirichlet): `iters` gradient-descent iterations with
    Armijo backtracking of the nodal log D field, from log D = log 0.1 everywhere, to
    match the state at t = tf of the workload's own field (log D ~ N(log 0.1, 0.3^2)).
    Forward: nsteps RKC steps (damping 10, stages covering 1.25 dt rho(J) of the true
    field); gradient: their discrete adjoint on fvm_vjp."""
    from finitevolumemethod import adjoint as AD
    from finitevolumemethod import workloads as WL

    tri, prob, u, logd_true, _ = WL.differentiation_problem(n_side, final_time=tf)
    t0 = time.time()
    fit = AD.DiffusivityFit(prob, tf / nsteps, nsteps, method="rkc")
    target = fit.forward(logd_true)
    t_setup = time.time() - t0
    logd0 = np.full(len(u), np.log(0.1))
    counts = {"rhs": 0, "vjp": 0}
    ctx = fit.ctx
    rhs0, vjp0 = ctx.rhs, ctx.vjp

    def rhs(*a, **k):
        counts["rhs"] += 1
        return rhs0(*a, **k)

    def vjp(*a, **k):
        counts["vjp"] += 1
        return vjp0(*a, **k)

    ctx.rhs, ctx.vjp = rhs, vjp
    try:
        t0 = time.time()
        logd, hist = AD.fit_log_diffusivity(prob, target, logd0, None, None, iters=iters, fit=fit)
        sec = time.time() - t0
    finally:
        ctx.rhs, ctx.vjp = rhs0, vjp0
    err0 = float(np.linalg.norm(logd0 - logd_true) / np.linalg.norm(logd_true))
    err1 = float(np.linalg.norm(logd - logd_true) / np.linalg.norm(logd_true))
    out = dict(workload=f"log D fit, {n_side}x{n_side} jittered mesh, u0~U[0,1), target u(t={tf:g}) of "
                        "log D~N(log 0.1,0.3^2), start log D=log 0.1, gradient descent with Armijo backtracking",
               nodes=int(tri.num_points()), iterations=len(hist), rkc_steps=nsteps, rkc_stages=fit.stages,
               rkc_damping=10.0, dt=tf / nsteps, setup_seconds=round(t_setup, 2),
               loss_first=hist[0], loss_last=hist[-1], loss_drop=round(hist[-1] / hist[0], 6),
               seconds=round(sec, 2), seconds_per_iteration=round(sec / len(hist), 3),
               rhs_evaluations=counts["rhs"], vjp_evaluations=counts["vjp"],
               logd_rel_error_start=round(err0, 4), logd_rel_error_end=round(err1, 4))
    log(f"fit {out}")
    return out


def _solve_all(dt, rtol, tag):
    """Each device solver on one template operator (after one warm-up solve), CG with
    its direction kernel on the ELL copy of M (the default) and, for comparison, on
    CSR: (per-solver stats, host solutions)."""
    from finitevolumemethod import templates as TM

    out, sols = {}, {}
    for name, alg in (("cg_fp64", TM.CG()), ("mixed_fp32_fp64", TM.MixedCG()),
                      ("mixed_restart_fp32_fp64", TM.MixedCG(update="restart")),
                      ("bicgstab_fp64", TM.BiCGStab())):
        dt.solve(alg, rtol=rtol)
        x, st = dt.solve(alg, rtol=rtol)
        sols[name] = x.to_host()
        out[name] = dict(seconds=round(st["seconds"], 4), iterations=st["iterations"],
                         outer=st["outer"], residual=st["residual"], converged=st["converged"])
        log(f"template {tag} {name}: {st}")
    ref = sols["cg_fp64"]
    out["max_rel_diff_vs_cg"] = {k: float(np.linalg.norm(v - ref) / np.linalg.norm(ref))
                                 for k, v in sols.items() if k != "cg_fp64"}
    out["mixed_speedup_vs_cg"] = round(out["cg_fp64"]["seconds"] / out["mixed_fp32_fp64"]["seconds"], 3)
    # the CG direction kernel on the CSR form of M instead of its ELL copy
    csr = {}
    for name, alg in (("cg_fp64", TM.CG()), ("mixed_fp32_fp64", TM.MixedCG())):
        dt.solve(alg, rtol=rtol, spmv_format="csr")
        x, st = dt.solve(alg, rtol=rtol, spmv_format="csr")
        csr[name] = dict(seconds=round(st["seconds"], 4), iterations=st["iterations"], converged=st["converged"],
                         ell_speedup=round(st["seconds"] / out[name]["seconds"], 3))
        log(f"template {tag} {name} (CSR form): {st}")
    csr["mixed_speedup_vs_cg"] = round(csr["cg_fp64"]["seconds"] / csr["mixed_fp32_fp64"]["seconds"], 3)
    out["csr_form"] = csr
    return out, sols


def template_solve(n_side, rtol=1e-10, direct=True):
    """The steady template workload: PoissonsEquation and MeanExitTimeProblem on an
    n_side x n_side jittered mesh of [0, 1]^2, homogeneous Dirichlet,
    D = exp(0.5 sin 4 pi x cos 4 pi y), f ~ U[-1, 1] (Poisson).  Time to
    |b - A x| <= rtol |b| for each device solver (after one warm-up solve), the SpMV
    rates of the solver operator, and (direct=True) each solution's relative error
    against an FP64 sparse LU solve (SuperLU, COLAMD ordering) of the same A and b on
    the host (build_seconds: both templates, mesh included)."""
    from finitevolumemethod import _lib
    from finitevolumemethod import workloads as WL

    t0 = time.time()
    tri, prob, met, f = WL.template_problems(n_side)
    t_build = time.time() - t0
    N = tri.num_points()
    dt = prob.problem
    out = dict(workload=f"PoissonsEquation and MeanExitTimeProblem {n_side}x{n_side} jittered, "
                        "D=exp(0.5 sin4pix cos4piy), f~U[-1,1] (Poisson), homogeneous Dirichlet",
               nodes=int(N), nnz=int(dt.nnz), rtol=rtol, build_seconds=round(t_build, 3))
    res, sols = _solve_all(dt, rtol, "poisson")
    out.update(res)
    out["mean_exit_time"], met_sols = _solve_all(met.problem, rtol, "mean_exit_time")
    x = _lib.DeviceArray.from_host(sols["cg_fp64"])
    y = _lib.DeviceArray(N)
    W, n_over, ib = dt.ell_info(index_bytes=True)
    n_off = int(dt.nnz) - N
    n_slices = (N + 63) // 64
    out["operator_form"] = dict(layout="sliced ELL: 64-row slices (one wave), rows ordered by length inside "
                                       "a slice, slot s holding the s-th off-diagonal of the rows that have one",
                                unrolled_slots=W, off_diagonal_entries=n_off, entries_past_unrolled_slots=n_over,
                                column_index_bytes=ib)
    # bytes per product on the sliced ELL copy: each off-diagonal's (column, value),
    # per row the diagonal, the lane's row byte, x gathered once and y written once,
    # per slice its W + 1 slot starts and remainder pointer
    for which, key, s in ((1, "spmv_fp64", 8), (2, "spmv_fp32", 4)):
        nbytes = n_off * (ib + s) + N * (s + 1 + 8 + 8) + n_slices * 4 * (W + 2)
        for _ in range(3):
            dt.spmv(which, x, y)
        tm = _lib.Timer()
        tm.start()
        reps = 50
        for _ in range(reps):
            dt.spmv(which, x, y)
        ms = tm.stop() / reps
        out[key] = dict(us=round(ms * 1e3, 2), GBps=round(nbytes / (ms * 1e-3) / 1e9, 1))
    if direct:
        import scipy.sparse.linalg as spla

        t0 = time.time()
        A = prob.A
        lu = spla.splu(A.tocsc(), permc_spec="COLAMD")
        A2 = met.A
        same = (A.shape == A2.shape and np.array_equal(A.indptr, A2.indptr)
                and np.array_equal(A.indices, A2.indices) and np.array_equal(A.data, A2.data))
        lu2 = lu if same else spla.splu(A2.tocsc(), permc_spec="COLAMD")
        errs = {}
        for tag, M, b, fac, ss in (("poisson", A, prob.b, lu, sols),
                                   ("mean_exit_time", A2, met.b, lu2, met_sols)):
            xd = fac.solve(b)
            nd = np.linalg.norm(xd)
            errs[tag] = dict(direct_residual=float(np.linalg.norm(M @ xd - b) / np.linalg.norm(b)),
                             **{k: float(np.linalg.norm(v - xd) / nd) for k, v in ss.items()})
        out["error_vs_fp64_direct"] = dict(
            solver="scipy SuperLU (splu, COLAMD), FP64, on the host copy of the device-assembled A, b",
            seconds=round(time.time() - t0, 1), shared_factorisation=bool(same), **errs)
        log(f"direct FP64 check {out['error_vs_fp64_direct']}")
    return out


def solve_bench(prob, rhs_ms, tf=4e-6):
    """End-to-end explicit integration of the throughput problem: Tsit5 (abstol 1e-6,
    reltol 1e-3, the library's integrator: every stage, the error norm, the step
    control and the Dirichlet update on the device) from t = 0 to tf, after one
    warm-up solve.  The i.i.d. random state makes the steps stability-bound (about
    1e-7), so tf fixes roughly 40 steps.  `seconds`: solve() with the state copied in
    and the saved states copied out; `seconds_device_resident`: fvm_ode_solve on a
    state already on the device, no saves (ms_per_step and rhs_share use this)."""
    import finitevolumemethod as FVM

    keep = prob.final_time
    try:
        prob.final_time = tf / 8
        FVM.solve(prob, FVM.Tsit5(), saveat=[0.0, prob.final_time])
        prob.final_time = tf
        t0 = time.perf_counter()
        sol = FVM.solve(prob, FVM.Tsit5(), saveat=[0.0, tf])
        sec = time.perf_counter() - t0
    finally:
        prob.final_time = keep
    # the integration alone: the state already on the device, no saved states
    from finitevolumemethod import _lib
    from finitevolumemethod.solve import fsal_ok

    ctx = prob.device_context()
    d_u = _lib.DeviceArray.from_host(prob.initial_condition)
    o = _lib.OdeOpts()
    o.method, o.op_kind, o.op, o.dirichlet = _lib.ODE_TSIT5, _lib.OP_PROBLEM, ctx.handle, 1
    o.fsal = int(fsal_ok(prob))  # what solve() passes: homogeneous Dirichlet values are constant
    o.n, o.t0, o.tf, o.dt = len(prob.initial_condition), 0.0, tf, 0.0
    o.abstol, o.reltol, o.maxiters, o.n_save = 1e-6, 1e-3, 10**7, 0
    ost = _lib.OdeStats()
    t0 = time.perf_counter()
    _lib.check(_lib.hip_lib().fvm_ode_solve(ctypes.byref(o), d_u.ptr, ctypes.byref(ost), None), "fvm_ode_solve")
    dev = time.perf_counter() - t0
    st = sol.stats
    out = dict(workload=f"Tsit5 on the throughput problem from t = 0 to {tf:g} (abstol 1e-6, reltol 1e-3)",
               retcode=sol.retcode, seconds=round(sec, 4), seconds_device_resident=round(dev, 4),
               accepted_steps=int(st["naccept"]), rejected_steps=int(st["nreject"]),
               rhs_evaluations=int(st["nf"]), ms_per_step=round(1e3 * dev / max(1, ost.naccept), 4),
               rhs_share=round(ost.nf * rhs_ms * 1e-3 / dev, 3))
    log(f"solve {out}")
    return out


def published_walkthroughs():
    """The reference's own @btime'd template walkthroughs, run end to end through solve()
    (best of 3 after one warm-up), beside the published times (their CPU is unstated:
    context, not the metric).  diffusion_equations.jl:327-331 (50 x 50 plate on [0, 2]^2,
    D = 1/9, Dirichlet 0, Tsit5, saveat 0.05: 5.736 ms) and :343-370, 442-446 (100 x 100
    on [0, 320]^2, D = 2, Neumann, u0 = 1 on 0.4 L <= y <= 0.6 L, Tsit5 to 500, saveat 100:
    78.761 ms), with the FVMProblem forms of both (:336-340, TRBDF2 with KLU: 49.237 ms;
    :451-455, CVODE_BDF with GMRES: 94.839 ms); poissons_equation.jl:92-95, 123-130, 168-180 (laplace u = -sin(pi x) sin(pi y) on
    a 100 x 100 grid of [0, 1]^2, u = 0 on the boundary, as the PoissonsEquation template:
    15.442 ms with KLU, and as a SteadyFVMProblem: 189.406 ms); laplaces_equation.jl:161-222
    (D = (x + 1)(y + 2) on a 100 x 100 grid of [0, 5]^2: 15.368 ms with KLU, 495.417 ms as a
    SteadyFVMProblem); linear_reaction_diffusion_equations.jl:82-186 (150 x 150 grid of
    [0, 1]^2 to t = 8, saveat 2: 48.360 ms for the template and 163.686 ms for the
    FVMProblem, both with CVODE_BDF and GMRES); poissons_equation.jl:185-382 (the
    parallel-plate capacitor: 9.061 ms with KLU, 329.327 ms as a SteadyFVMProblem), on a
    grid of comparable size instead of the reference's Delaunay-refined mesh;
    mean_exit_time.jl:89-232 (the perturbed annulus: 2.559 ms with KLU, 221.851 ms as a
    SteadyFVMProblem with KenCarp47), on a ring mesh of the same domain."""
    import finitevolumemethod as FVM

    def best(fn, stats=False):
        fn()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            sol = fn()
            ts.append(time.perf_counter() - t0)
        ms = round(1e3 * min(ts), 2)
        if not stats:
            return ms
        keep = ("naccept", "nreject", "nnewton", "nkrylov", "device_loop", "jacobian_colours")
        return ms, {k: v for k, v in getattr(sol, "stats", {}).items() if k in keep}

    out = []
    tri = FVM.triangulate_rectangle(0.0, 2.0, 0.0, 2.0, 50, 50, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    plate = FVM.DiffusionEquation(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0, FVM.Dirichlet),
                                  diffusion_function=lambda x, y, p: 1 / 9,
                                  initial_condition=np.where(tri.y <= 1.0, 50.0, 0.0), final_time=0.5)
    out.append(dict(workload="DiffusionEquation 50x50 plate, Tsit5, saveat 0.05", ms=best(
        lambda: FVM.solve(plate, FVM.Tsit5(), saveat=0.05)), published_ms=5.736,
        source="docs/src/literate_wyos/diffusion_equations.jl:327-331"))
    plate_fvm = FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet),
                               diffusion_function=lambda x, y, t, u, p: 1 / 9 + 0.0 * u,
                               initial_condition=np.where(tri.y <= 1.0, 50.0, 0.0), final_time=0.5)
    def first_steady(make):
        # a fresh SteadyFVMProblem: its device context, Jacobian extraction and linear set-up in the time
        prob = make()
        t0 = time.perf_counter()
        FVM.solve(prob)
        return round(1e3 * (time.perf_counter() - t0), 2)

    def first(make, saveat):
        # a fresh problem object: its device context and Jacobian extraction in the time
        prob = make()
        t0 = time.perf_counter()
        FVM.solve(prob, FVM.TRBDF2(), saveat=saveat)
        return round(1e3 * (time.perf_counter() - t0), 2)

    def make_plate():
        return FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet),
                              diffusion_function=lambda x, y, t, u, p: 1 / 9 + 0.0 * u,
                              initial_condition=np.where(tri.y <= 1.0, 50.0, 0.0), final_time=0.5)

    ms, st = best(lambda: FVM.solve(plate_fvm, FVM.TRBDF2(), saveat=0.05), stats=True)
    out.append(dict(workload="the plate as an FVMProblem, TRBDF2 (ours: the affine problem's Jacobian from "
                             "coloured fvm_jvp probes, then the whole integration with Newton-GMRES in one "
                             "workgroup; published: KLU on the sparse Jacobian), saveat 0.05",
                    ms=ms, first_solve_ms=first(make_plate, 0.05), published_ms=49.237, stats=st,
                    note="ms: repeated solves of one problem (the extracted Jacobian is kept on it); "
                         "first_solve_ms: a fresh problem, its device set-up and Jacobian extraction included",
                    source="docs/src/literate_wyos/diffusion_equations.jl:336-340"))
    L = 320.0
    tri = FVM.triangulate_rectangle(0.0, L, 0.0, L, 100, 100, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    big = FVM.DiffusionEquation(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 2.0, FVM.Neumann),
                                diffusion_function=lambda x, y, p: 2.0,
                                initial_condition=np.where((0.4 * L <= tri.y) & (tri.y <= 0.6 * L), 1.0, 0.0),
                                final_time=500.0)
    out.append(dict(workload="DiffusionEquation 100x100 on [0,320]^2, Neumann, Tsit5, saveat 100", ms=best(
        lambda: FVM.solve(big, FVM.Tsit5(), saveat=100.0)), published_ms=78.761,
        source="docs/src/literate_wyos/diffusion_equations.jl:442-446"))
    big_fvm = FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 2.0 + 0.0 * u, FVM.Neumann),
                             diffusion_function=lambda x, y, t, u, p: 2.0 + 0.0 * u,
                             initial_condition=np.where((0.4 * L <= tri.y) & (tri.y <= 0.6 * L), 1.0, 0.0),
                             final_time=500.0)
    def make_big():
        return FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 2.0 + 0.0 * u, FVM.Neumann),
                              diffusion_function=lambda x, y, t, u, p: 2.0 + 0.0 * u,
                              initial_condition=np.where((0.4 * L <= tri.y) & (tri.y <= 0.6 * L), 1.0, 0.0),
                              final_time=500.0)

    ms, st = best(lambda: FVM.solve(big_fvm, FVM.TRBDF2(), saveat=100.0), stats=True)
    out.append(dict(workload="the same as an FVMProblem, stiff (ours: TRBDF2 in one workgroup as above; "
                             "published: CVODE_BDF with GMRES), saveat 100",
                    ms=ms, first_solve_ms=first(make_big, 100.0), published_ms=94.839, stats=st,
                    source="docs/src/literate_wyos/diffusion_equations.jl:451-455"))
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 100, 100, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * x, FVM.Dirichlet)
    src = lambda x, y, p: -np.sin(np.pi * x) * np.sin(np.pi * y)  # noqa: E731
    pois = FVM.PoissonsEquation(mesh, BCs, source_function=src)
    out.append(dict(workload="PoissonsEquation 100x100 on [0,1]^2, device CG to 1e-10 (published: KLU)", ms=best(
        lambda: FVM.solve(pois)), published_ms=15.442,
        source="docs/src/literate_wyos/poissons_equation.jl:168-172"))
    def make_steady():
        return FVM.SteadyFVMProblem(FVM.FVMProblem(
            mesh, BCs, diffusion_function=lambda x, y, t, u, p: 1.0 + 0.0 * u,
            source_function=lambda x, y, t, u, p: -src(x, y, None) + 0.0 * u,
            initial_condition=np.zeros(tri.num_points()), final_time=np.inf))

    steady = make_steady()
    out.append(dict(workload="the same Poisson problem as a SteadyFVMProblem (ours: fvm_eqs! is affine in u, so one "
                             "device linear solve on its Jacobian from coloured fvm_jvp probes, Newton-GMRES "
                             "otherwise; published: DynamicSS(TRBDF2(KLU)))", ms=best(lambda: FVM.solve(steady)),
                    first_solve_ms=first_steady(make_steady), published_ms=189.406,
                    note="ms: repeated solves of one problem (its extracted Jacobian and linear set-up are kept); "
                         "first_solve_ms: a fresh problem, those included",
                    source="docs/src/literate_wyos/poissons_equation.jl:176-180"))
    # poissons_equation.jl:185-382: the parallel-plate capacitor on [0, 10]^2, plates at
    # 2 <= x <= 8 on y = 3 (V = -1) and y = 7 (V = 1) as internal Dirichlet nodes, eps(x) =
    # 1 + (eps0 - 1)(1 + erf(r / 4)) / 2, source -Q exp(-r^2 / 2) / (2 pi eps0), r the distance
    # to the plates; Dirichlet 0 at the bottom and top, zero flux at the sides.  The reference
    # meshes it by Delaunay refinement to max_area = 1e-4 of the area (DelaunayTriangulation,
    # absent here); this is a 101 x 101 grid (triangles of area 0.005 against the bound 0.01)
    # on which the plates are grid lines.
    from scipy.special import erf

    def plate_dist(x, y):
        x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
        dx = x - np.clip(x, 2.0, 8.0)
        return np.minimum(np.hypot(dx, y - 3.0), np.hypot(dx, y - 7.0))

    eps0, Q = 8.8541878128e-12, 1.0e-6
    diel = lambda x, y, p: 1 + (eps0 - 1) * (1 + erf(plate_dist(x, y) / 4.0)) / 2  # noqa: E731
    plate_src = lambda x, y, p: -(Q / (2 * np.pi) * np.exp(-plate_dist(x, y) ** 2 / 2)) / eps0  # noqa: E731
    tri = FVM.triangulate_rectangle(0.0, 10.0, 0.0, 10.0, 101, 101, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    zero_f = lambda x, y, t, u, p: 0.0 * x  # noqa: E731
    BCs = FVM.BoundaryConditions(mesh, (zero_f, zero_f, zero_f, zero_f),
                                 (FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet, FVM.Neumann))
    on = (tri.x >= 2.0 - 1e-9) & (tri.x <= 8.0 + 1e-9)
    lower = np.nonzero(on & (np.abs(tri.y - 3.0) < 1e-9))[0]
    upper = np.nonzero(on & (np.abs(tri.y - 7.0) < 1e-9))[0]
    ICs = FVM.InternalConditions((lambda x, y, t, u, p: -1.0 + 0.0 * x, lambda x, y, t, u, p: 1.0 + 0.0 * x),
                                 dirichlet_nodes={**{int(i): 0 for i in lower}, **{int(i): 1 for i in upper}})
    cap = FVM.PoissonsEquation(mesh, BCs, ICs, diffusion_function=diel, source_function=plate_src)
    out.append(dict(workload="PoissonsEquation, parallel-plate capacitor on [0,10]^2 (101x101 grid; the reference "
                             "refines a Delaunay mesh to max_area 1e-4 of the domain), device CG to 1e-10 "
                             "(published: KLU)", ms=best(lambda: FVM.solve(cap)), published_ms=9.061,
                    mesh_note="comparable mesh, not the reference's", nodes=int(tri.num_points()),
                    source="docs/src/literate_wyos/poissons_equation.jl:207-374"))
    def make_cap():
        return FVM.SteadyFVMProblem(FVM.FVMProblem(
            mesh, BCs, ICs, diffusion_function=lambda x, y, t, u, p: diel(x, y, None) + 0.0 * u,
            source_function=lambda x, y, t, u, p: -plate_src(x, y, None) + 0.0 * u,
            initial_condition=np.zeros(tri.num_points()), final_time=np.inf))

    cap_fvm = make_cap()
    ms = best(lambda: FVM.solve(cap_fvm))
    st = FVM.solve(cap_fvm).stats
    out.append(dict(workload="the same as a SteadyFVMProblem (ours: one device linear solve on the Jacobian; "
                             "published: DynamicSS(TRBDF2(KLU)))", ms=ms, first_solve_ms=first_steady(make_cap),
                    published_ms=329.327,
                    stats=dict(newton_iters=st["newton_iters"], linear_solve=st.get("linear_solve")),
                    mesh_note="comparable mesh, not the reference's",
                    source="docs/src/literate_wyos/poissons_equation.jl:352-382"))
    # mean_exit_time.jl:89-232: the annulus 1 <= r <= 3 with a perturbed interface between
    # D = 6.25e-4 and 6.25e-5, T = 0 on a short outer arc, the hole and two point holes, zero
    # flux on the rest (workloads.met_problems: rings and the interface points, Delaunay,
    # instead of the reference's constrained refinement to max_area 1e-3 of the area)
    from finitevolumemethod import workloads as WL

    tri_m, met, met_steady = WL.met_problems()
    out.append(dict(workload="MeanExitTimeProblem on the perturbed annulus with point holes (2813-point ring mesh; "
                             "the reference refines a constrained Delaunay mesh), device CG to 1e-10 "
                             "(published: KLU)", ms=best(lambda: FVM.solve(met)), published_ms=2.559,
                    mesh_note="comparable mesh, not the reference's", nodes=int(tri_m.num_points()),
                    source="docs/src/literate_wyos/mean_exit_time.jl:89-224"))
    ms = best(lambda: FVM.solve(met_steady))
    st = FVM.solve(met_steady).stats
    out.append(dict(workload="the same as a SteadyFVMProblem (ours: one device linear solve on the Jacobian; "
                             "published: DynamicSS(KenCarp47(KLU)))", ms=ms,
                    first_solve_ms=first_steady(lambda: WL.met_problems()[2]), published_ms=221.851,
                    stats=dict(newton_iters=st["newton_iters"], linear_solve=st.get("linear_solve")),
                    mesh_note="comparable mesh, not the reference's",
                    source="docs/src/literate_wyos/mean_exit_time.jl:160-232"))
    # laplaces_equation.jl:161-222: D = (x + 1)(y + 2) on [0, 5]^2, u = 0 / 5 on x = 0 / 5,
    # zero flux on y = 0, 5
    tri = FVM.triangulate_rectangle(0.0, 5.0, 0.0, 5.0, 100, 100, single_boundary=False)
    mesh = FVM.FVMGeometry(tri)
    zero_f = lambda x, y, t, u, p: 0.0 * x  # noqa: E731
    five_f = lambda x, y, t, u, p: 5.0 + 0.0 * x  # noqa: E731
    BCs = FVM.BoundaryConditions(mesh, (zero_f, five_f, zero_f, zero_f),
                                 (FVM.Neumann, FVM.Dirichlet, FVM.Neumann, FVM.Dirichlet))
    lap = FVM.LaplacesEquation(mesh, BCs, diffusion_function=lambda x, y, p: (x + 1) * (y + 2))
    out.append(dict(workload="LaplacesEquation 100x100 on [0,5]^2, D=(x+1)(y+2), Dirichlet / Neumann sides, "
                             "device CG to 1e-10 (published: KLU)", ms=best(lambda: FVM.solve(lap)),
                    published_ms=15.368, source="docs/src/literate_wyos/laplaces_equation.jl:161-215"))
    def make_lap():
        return FVM.SteadyFVMProblem(FVM.FVMProblem(
            mesh, BCs, diffusion_function=lambda x, y, t, u, p: (x + 1) * (y + 2) + 0.0 * u,
            initial_condition=np.zeros(tri.num_points()), final_time=np.inf))

    lap_fvm = make_lap()
    ms = best(lambda: FVM.solve(lap_fvm))
    st = FVM.solve(lap_fvm).stats
    ms_newton = best(lambda: FVM.solve(lap_fvm, linear="newton"))
    st_newton = FVM.solve(lap_fvm, linear="newton").stats
    out.append(dict(workload="the same as a SteadyFVMProblem (ours: one device linear solve on the Jacobian as "
                             "above; published: DynamicSS(TRBDF2(KLU)))", ms=ms, first_solve_ms=first_steady(make_lap),
                    published_ms=495.417,
                    stats=dict(newton_iters=st["newton_iters"], linear_solve=st.get("linear_solve")),
                    newton_gmres=dict(ms=ms_newton, newton_iters=st_newton["newton_iters"],
                                      nkrylov=st_newton["nkrylov"], precond=st_newton["precond"]),
                    source="docs/src/literate_wyos/laplaces_equation.jl:194-222"))
    # linear_reaction_diffusion_equations.jl:82-184: dT/dt = div[1e-3 x^2 y grad T] + (x-1)(y-1) T
    # on a 150 x 150 grid of [0, 1]^2, grad T . n = 1, T0 = x^2 + y^2, to t = 8, saveat 2
    tri = FVM.triangulate_rectangle(0.0, 1.0, 0.0, 1.0, 150, 150, single_boundary=True)
    mesh = FVM.FVMGeometry(tri)
    Dfun = lambda x, y, p: 1.0e-3 * x ** 2 * y  # noqa: E731
    Sfun = lambda x, y, p: (x - 1) * (y - 1)  # noqa: E731
    ic = tri.x ** 2 + tri.y ** 2
    lrd = FVM.LinearReactionDiffusionEquation(
        mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 1.0 + 0.0 * x, FVM.Neumann),
        diffusion_function=Dfun, source_function=Sfun, initial_condition=ic, final_time=8.0)
    out.append(dict(workload="LinearReactionDiffusionEquation 150x150 on [0,1]^2, D=1e-3 x^2 y, f=(x-1)(y-1), "
                             "Neumann, to t=8, saveat 2 (ours: Tsit5, the tutorial's own solve; published: "
                             "CVODE_BDF with GMRES)", ms=best(lambda: FVM.solve(lrd, FVM.Tsit5(), saveat=2.0)),
                    published_ms=48.360, source="docs/src/literate_wyos/linear_reaction_diffusion_equations.jl:172-178"))

    def make_lrd():
        # the Neumann condition as q . n = -D (x, y) (:119-123)
        return FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: -Dfun(x, y, None) + 0.0 * u,
                                                           FVM.Neumann),
                              diffusion_function=lambda x, y, t, u, p: Dfun(x, y, None) + 0.0 * u,
                              source_function=lambda x, y, t, u, p: Sfun(x, y, None) * u,
                              initial_condition=ic, final_time=8.0)

    lrd_fvm = make_lrd()
    ms, st = best(lambda: FVM.solve(lrd_fvm, FVM.TRBDF2(), saveat=2.0), stats=True)
    out.append(dict(workload="the same as an FVMProblem, stiff (ours: TRBDF2 in one workgroup; published: "
                             "CVODE_BDF with GMRES), saveat 2", ms=ms, first_solve_ms=first(make_lrd, 2.0),
                    published_ms=163.686, stats=st,
                    source="docs/src/literate_wyos/linear_reaction_diffusion_equations.jl:180-186"))
    log(f"published walkthroughs {out}")
    return out


def ensemble_bench(P=256, steps=1000, graph_steps=100):
    """The ensemble workload (workloads.ensemble_members): P rectangles [0, 2]^2 with
    nx, ny ~ round(exp(U[log 20, log 300])) drawn independently, interior points
    jittered by U(-h/4, h/4), constant D_p ~ LogNormal(log 0.1, 0.5), homogeneous
    Dirichlet, u0 = 50 where y <= 1 (mid-height), else 0, forward Euler with dt_p =
    0.9 x the Gershgorin stability limit, `steps` steps.  Times the packed ensemble
    (autotuned tiles, graph replay) against looping the single-problem path over the
    members (each member's steps issued back to back from C)."""
    from finitevolumemethod import _lib
    from finitevolumemethod import ensemble as EN
    from finitevolumemethod import workloads as WL

    t0 = time.time()
    probs, _ = WL.ensemble_members(P)
    dts = [EN.explicit_dt(q) for q in probs]
    N = sum(q.mesh.triangulation.num_points() for q in probs)
    T = sum(len(q.mesh.triangulation.triangles) for q in probs)
    log(f"ensemble: {P} members, {N} nodes, {T} triangles ({time.time() - t0:.1f}s setup)")
    t0 = time.time()
    b = EN.DeviceBatch(probs, dts, autotune=True, graph_steps=graph_steps)
    log(f"packed batch built and autotuned in {time.time() - t0:.1f}s: tile {b.info.tile_triangles}, "
        f"candidates {list(b.info.candidate_tiles)[:b.info.n_candidates]} "
        f"{[round(v, 4) for v in list(b.info.candidate_ms)[:b.info.n_candidates]]} ms")
    u0 = b.pack([q.initial_condition for q in probs])
    d_u = _lib.DeviceArray.from_host(u0)
    b.euler(d_u, graph_steps, 0.0)  # warm-up (capture + one replay)
    d_u.upload(u0)
    _lib.synchronize()
    tm = _lib.Timer()
    tm.start()
    b.euler(d_u, steps, 0.0)
    ms_batch = tm.stop()
    batched = d_u.to_host()
    # the same steps without the graph (one launch sequence per step)
    b0 = EN.DeviceBatch(probs, dts, tile_triangles=b.info.tile_triangles, graph_steps=0)
    d_u.upload(u0)
    b0.euler(d_u, 10, 0.0)
    d_u.upload(u0)
    _lib.synchronize()
    tm.start()
    b0.euler(d_u, steps, 0.0)
    ms_nograph = tm.stop()
    # the update as a separate kernel after the RHS (graph replay as above)
    bu = EN.DeviceBatch(probs, dts, tile_triangles=b.info.tile_triangles, graph_steps=graph_steps, fused=False)
    d_u.upload(u0)
    bu.euler(d_u, graph_steps, 0.0)
    d_u.upload(u0)
    _lib.synchronize()
    tm.start()
    bu.euler(d_u, steps, 0.0)
    ms_unfused = tm.stop()
    del bu
    # looping the single-problem path over the members
    singles = [EN.DeviceBatch([q], [dt], graph_steps=0) for q, dt in zip(probs, dts)]
    us = [_lib.DeviceArray.from_host(q.initial_condition) for q in probs]
    for s, u in zip(singles, us):
        s.euler(u, 5, 0.0)
    for q, u in zip(probs, us):
        u.upload(q.initial_condition)
    _lib.synchronize()
    tm.start()
    for s, u in zip(singles, us):
        s.euler(u, steps, 0.0)
    ms_loop = tm.stop()
    looped = np.concatenate([u.to_host() for u in us])
    err = float(np.max(np.abs(batched - looped)) / max(1e-300, np.max(np.abs(looped))))
    per_step = 132 * T + 25 * N
    out = dict(workload=f"{P} jittered rectangles on [0,2]^2, nx,ny~logU[20,300], D~LogNormal(log 0.1,0.5), "
                        f"u0=50 for y<=1, homogeneous Dirichlet, forward Euler dt=0.9x stability limit, "
                        f"{steps} steps",
               members=P, nodes=int(N), triangles=int(T), steps=steps,
               tile_triangles=int(b.info.tile_triangles),
               autotune_ms={int(t): round(float(m), 5) for t, m in
                            zip(list(b.info.candidate_tiles)[:b.info.n_candidates],
                                list(b.info.candidate_ms)[:b.info.n_candidates])},
               launches_per_step=2,
               launches_note="tile kernel + shared-node fix-up, the Euler update and the Dirichlet "
                             "callback fused into both; one graph replay holds graph_steps steps",
               graph_steps=graph_steps, us_per_step=round(1e3 * ms_batch / steps, 2),
               batched_graph_ms=round(ms_batch, 3), batched_nograph_ms=round(ms_nograph, 3),
               batched_unfused_update_ms=round(ms_unfused, 3),
               looped_ms=round(ms_loop, 3), speedup_vs_loop=round(ms_loop / ms_batch, 2),
               algorithmic_GBps=round(per_step * steps / (ms_batch * 1e-3) / 1e9, 1),
               max_rel_diff_vs_loop=err)
    log(f"ensemble: batched {ms_batch:.1f} ms (no graph {ms_nograph:.1f} ms, unfused update {ms_unfused:.1f} ms), "
        f"looped {ms_loop:.1f} ms, "
        f"diff {err:.2e}")
    return out


def system_bench(mesh, reps=50):
    """FVMSystem RHS (fvm_system_rhs) of the Gray-Scott tutorial's closures (two
    variables, q_v = -eps_v grad u_v, S = b (1 - u) - u v^2 and -d v + u v^2,
    zero-flux boundary) on the throughput mesh, u, v ~ U[0, 1).  Canonical bytes per
    RHS: the 204-B triangle record, per node 8 nv (u) + 8 (V) + 8 nv (du) + nv (codes),
    44 B per boundary edge."""
    import finitevolumemethod as FVM
    from finitevolumemethod import _lib

    tri = mesh.triangulation
    N, T, B = tri.num_points(), len(tri.triangles), len(tri.boundary_edges)
    rng = np.random.default_rng(SEED + 11)
    bc = lambda x, y, t, u, p: 0.0 * u[0] * 0.0 * u[1]  # noqa: E731
    probs = [FVM.FVMProblem(mesh, FVM.BoundaryConditions(mesh, bc, FVM.Neumann),
                            flux_function=lambda x, y, t, a, b, g, e, v=v: (-a[v] * e, -b[v] * e),
                            flux_parameters=eps, source_function=S, source_parameters=sp,
                            initial_condition=rng.random(N), final_time=1.0)
             for v, eps, S, sp in ((0, 2e-5, lambda x, y, t, u, b: b * (1 - u[0]) - u[0] * u[1] ** 2, 0.04),
                                   (1, 1e-5, lambda x, y, t, u, d: -d * u[1] + u[0] * u[1] ** 2, 0.1))]
    sys_ = FVM.FVMSystem(*probs)
    ctx = sys_.device_context()
    d_u = _lib.DeviceArray.from_host(sys_.flat(sys_.initial_condition))
    d_du = _lib.DeviceArray(2 * N)
    for _ in range(5):
        ctx.rhs(d_u, d_du)
    tm = _lib.Timer()
    tm.start()
    for _ in range(reps):
        ctx.rhs(d_u, d_du)
    ms = tm.stop() / reps
    alg = canonical_bytes(N, T, B, per_node=8 * 2 + 8 + 8 * 2 + 2)
    return dict(workload="FVMSystem, Gray-Scott closures (2 variables), throughput mesh, zero-flux boundary",
                nodes=int(N), variables=2, rhs_ms=round(ms, 5),
                Gnode_updates_per_s=round(N / (ms * 1e-3) / 1e9, 3),
                canonical_bytes_per_rhs=int(alg),
                canonical_bytes_GBps=round(alg / (ms * 1e-3) / 1e9, 1),
                note="canonical bytes: the reference's data model (204-B triangle record); the kernel "
                     "rebuilds the record from coordinates, so these bytes are not moved")


def cpu_baseline(tri, u, budget_s, nthreads):
    """The oracle's restatement of parallel_fvm_eqs! (per-thread du columns, serial
    combine) on the same mesh and state, for about budget_s seconds."""
    from oracle import OracleProblem

    from finitevolumemethod.conditions import Conditions  # noqa: F401

    N = tri.num_points()
    kind = np.zeros(N, np.int8)
    fidx = np.full(N, -1, np.int32)
    bn = tri.boundary_nodes()
    kind[bn], fidx[bn] = 1, 0
    bneu = np.full(len(tri.boundary_edges), -1, np.int32)
    t0 = time.time()
    orc = OracleProblem(tri.x, tri.y, tri.triangles, tri.boundary_edges, tri.boundary_adjacent(),
                        kind, fidx, bneu,
                        dict(builtin=dict(flux_kind=0, dpoly=[0.1, 0, 1, 0], spoly=[0, 1, -1, 0],
                                          fn=([0.0], [0.0], [0.0]))))
    log(f"oracle geometry {time.time() - t0:.1f}s")
    orc.rhs(u, 0.0, nthreads=nthreads)  # warm-up
    reps, t_total = 0, 0.0
    while t_total < budget_s and reps < 200:
        t0 = time.perf_counter()
        du = orc.rhs(u, 0.0, nthreads=nthreads)
        t_total += time.perf_counter() - t0
        reps += 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return du, dict(value=round(N * reps / t_total / 1e9, 5), unit="Gnode-updates/s",
                    cores=nthreads, kind="port", cpu_model=model, nproc=usable,
                    cpu_count=os.cpu_count(),
                    sample=f"{reps} full evaluations of the same {N}-node mesh "
                           f"({t_total:.1f} s, OpenMP restatement of parallel_fvm_eqs!, "
                           f"{nthreads} threads)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n-side", type=int, default=2000)
    ap.add_argument("--vjp-side", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tile", type=int, default=0, help="triangles per tile (0: default 512)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--template-side", type=int, default=1000,
                    help="PoissonsEquation solve workload side (0: skip)")
    ap.add_argument("--no-direct", action="store_true",
                    help="skip the host sparse-LU error check of the template solves")
    ap.add_argument("--ensemble", type=int, default=256, help="ensemble members (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--fit-iters", type=int, default=50, help="log D fit iterations (0: skip)")
    ap.add_argument("--trbdf2-side", type=int, default=1000, help="TR-BDF2 workload side (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the finite-volume path runs on a single MI355X here "
              "(multi-GPU sharding is design-only); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import finitevolumemethod as FVM
    from finitevolumemethod import _lib

    _lib.require_gpu()
    tri, mesh, prob, u = build_throughput_problem(args.n_side, args.tile)
    ctx = prob.device_context()
    N, T, B = tri.num_points(), len(tri.triangles), len(tri.boundary_edges)
    d_u = _lib.DeviceArray.from_host(u)
    d_du = _lib.DeviceArray(N)
    for _ in range(args.warmup):
        ctx.rhs(d_u, d_du)
    _lib.synchronize()
    tm = _lib.Timer()
    tm.start()
    for _ in range(args.steps):
        ctx.rhs(d_u, d_du)
    ms_total = tm.stop()
    _lib.synchronize()
    ms_step = ms_total / args.steps
    value = N * args.steps / (ms_total * 1e-3) / 1e9
    # per-kernel durations (HIP events on the launch stream)
    tiles, fix = [], []
    for _ in range(max(5, min(50, args.steps))):
        a, b = rhs_timed(ctx, d_u, d_du)
        tiles.append(a)
        fix.append(b)
    ms_tiles, ms_fix = float(np.mean(tiles)), float(np.mean(fix))
    alg = canonical_bytes(N, T, B)
    # per-RHS counts from the committed rocprofv3 counter runs of this build and mesh:
    # HBM bytes (FETCH_SIZE + WRITE_SIZE, calibrated on the copy kernel) and VALU
    # instructions of the two RHS kernels
    cnt = None
    if os.path.exists(COUNTERS):
        c = json.load(open(COUNTERS))
        if (c.get("nodes") == N and c.get("tiles") == int(mesh.info.n_tiles)
                and c.get("build") == _lib.build_info()):
            cnt = c
    copy_gbps = copy_bandwidth()
    log(f"copy kernel {copy_gbps / 1e9:.0f} GB/s")
    du_gpu = d_du.to_host()
    log(f"RHS {ms_step * 1e3:.1f} us/step ({value:.3f} Gnode/s); tiles {ms_tiles * 1e3:.1f} us, "
        f"fix-up {ms_fix * 1e3:.1f} us (HIP events)")
    t_s = ms_step * 1e-3
    if cnt is not None:
        traffic = int(cnt["hbm_bytes_per_rhs"])
        valu = float(cnt["valu_insts_per_rhs"])  # wave64 VALU instructions, tile + fix-up
        achieved = valu * 64 * 2 / t_s  # FMA-equivalent flops of the issued VALU instructions
        roof = {"bound": "valu", "achieved": round(achieved / 1e12, 3), "peak": VALU_PEAK / 1e12,
                "unit": "TFLOP/s", "frac": round(achieved / VALU_PEAK, 4), "traffic": traffic,
                "valu_insts_per_rhs": int(valu),
                "fp64_flops_per_rhs": int(cnt["fp64_flops_per_rhs"]),
                "fp64_TFLOPs": round(cnt["fp64_flops_per_rhs"] / t_s / 1e12, 3),
                "valu_busy_tile_kernel_rocprof": cnt["tile_kernel"]["valu_busy"],
                "hbm": {"achieved": round(traffic / t_s / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": round(traffic / t_s / HBM_PEAK, 4),
                        "frac_of_copy_kernel": round(traffic / t_s / copy_gbps, 4)}}
        # the dominant kernel alone, over its own HIP-event duration
        tk = cnt["tile_kernel"]
        t_tk = ms_tiles * 1e-3
        roof["tile_kernel"] = {
            "ms": round(ms_tiles, 5), "valu_insts": int(tk["valu_insts"]),
            "achieved": round(tk["valu_insts"] * 128 / t_tk / 1e12, 3), "unit": "TFLOP/s",
            "frac": round(tk["valu_insts"] * 128 / t_tk / VALU_PEAK, 4),
            "hbm_GBps": round(tk["hbm_bytes"] / t_tk / 1e9, 1), "hbm_frac": round(tk["hbm_bytes"] / t_tk / HBM_PEAK, 4),
            "valu_busy_rocprof": tk["valu_busy"]}
    else:
        traffic = None
        roof = {"bound": "valu", "achieved": None, "peak": VALU_PEAK / 1e12, "unit": "TFLOP/s", "frac": None,
                "traffic": None, "note": "no counter profile of this build and mesh under profiles/"}
    roof.update({
        "basis": "ms_per_step (back-to-back RHS loop; tile + fix-up kernels and their launch gap)",
        "canonical_bytes_equivalent": {
            "bytes_per_rhs": int(alg), "GBps": round(alg / t_s / 1e9, 1),
            "note": "the reference's data model (204 B per triangle record, 25 B per node, 44 B per "
                    "boundary edge) over the step time; these bytes are not moved: the tile kernel "
                    "rebuilds each triangle record from the vertex coordinates"},
        "build": _lib.build_info(),
        "kernels": {"tile_kernel_ms": round(ms_tiles, 5), "fix_kernel_ms": round(ms_fix, 5),
                    "timing": "HIP events around each kernel on the launch stream"}})
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gnode-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fvm_eqs! nonlinear diffusion D(u)=0.1+u^2, S=u(1-u), "
                               f"{args.n_side}x{args.n_side} jittered Delaunay mesh on [0,1]^2, "
                               "homogeneous Dirichlet",
                   "nodes": N, "triangles": T, "boundary_edges": B,
                   "tiles": int(mesh.info.n_tiles), "fix_nodes": int(mesh.info.n_fix_nodes)},
        "roofline": roof,
        "copy_bandwidth_GBps": round(copy_gbps / 1e9, 1),
    }
    if not args.no_extras:
        out["vjp_rhs"] = vjp_ratio(args.vjp_side)
        log(f"VJP/RHS {out['vjp_rhs']}")
        if args.fit_iters > 0:
            out["log_d_fit"] = fit_bench(args.vjp_side, iters=args.fit_iters)
        out["solve_tsit5"] = solve_bench(prob, ms_step)
        if args.trbdf2_side > 0:
            out["solve_trbdf2"] = trbdf2_bench(args.trbdf2_side)
        out["published_walkthroughs"] = published_walkthroughs()
        out["system_rhs"] = system_bench(mesh)
        log(f"system RHS {out['system_rhs']}")
        if args.template_side > 0:
            out["template_solve"] = template_solve(args.template_side, direct=not args.no_direct)
        if args.ensemble > 0:
            out["ensemble"] = ensemble_bench(args.ensemble, args.ensemble_steps)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        du_cpu, cb = cpu_baseline(tri, u, args.cpu_seconds, threads)
        out["cpu_baseline"] = cb
        if "solve_tsit5" in out:  # the same integration's RHS evaluations alone, at the CPU rate
            sv = out["solve_tsit5"]
            sv["cpu_rhs_only_seconds"] = round(sv["rhs_evaluations"] * N / (cb["value"] * 1e9), 2)
            sv["speedup_vs_cpu_rhs_only"] = round(sv["cpu_rhs_only_seconds"] / sv["seconds_device_resident"], 1)
        err = float(np.max(np.abs(du_gpu - du_cpu)) / np.max(np.abs(du_cpu)))
        out["parity_max_rel_err_vs_cpu"] = err
        log(f"CPU baseline {cb['value']} Gnode/s on {threads} threads; max rel err {err:.2e}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
