"""solve(prob, alg; saveat, ...) with the state resident on the GPU
(reference src/solve.jl:133-220).

The reference hands fvm_eqs! to OrdinaryDiffEq.  This package carries its own
explicit integrators over device vectors (every stage is a fused linear
combination kernel plus one fvm_rhs launch) and applies the Dirichlet callback
(solve.jl:133-165, dirichlet.jl:78-86) after every accepted step, like the
reference's DiscreteCallback(true, update_dirichlet_nodes!).  Saved states at
`saveat` times come from cubic Hermite interpolation of the accepted steps.
"""
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .problem import FVMProblem, SteadyFVMProblem
from .system import FVMSystem

# Tsitouras (2011) 5(4) pair: nodes, stage weights (the 7th stage is the
# solution, first-same-as-last) and the error weights b - bhat.
_TSIT5_C = (0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0)
_TSIT5_A = (
    (),
    (0.161,),
    (-0.008480655492356989, 0.335480655492357),
    (2.897153057105493, -6.359448489975075, 4.3622954328695815),
    (5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525),
    (5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401,
     -0.028269050394068383),
    (0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081,
     2.324710524099774),
)
_TSIT5_BTILDE = (-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                 -0.1447110071732629, 0.5823571654525552, -0.45808210592918697,
                 0.015151515151515152)


@dataclass
class Tsit5:
    """Adaptive Tsitouras 5(4) with a PI step controller (beta1 = 7/50,
    beta2 = 2/25, safety 0.9, step ratio in [1/10, 5])."""
    pass


@dataclass
class Euler:
    """Forward Euler with a fixed step dt."""
    dt: float = None


@dataclass
class SSPRK3:
    """Shu-Osher three-stage strong-stability-preserving Runge-Kutta, fixed dt."""
    dt: float = None


@dataclass
class TRBDF2:
    """TR-BDF2 (Bank et al. 1985; Hosea & Shampine, Appl. Numer. Math. 20, 1996): a
    trapezoidal stage to t + gamma h, gamma = 2 - sqrt(2), then BDF2 through
    (t, t + gamma h, t + h); L-stable, order 2, the reference tutorials' stiff
    integrator.  Both stages solve z - d h f(z) = r, d = gamma / 2, by Newton with the
    Jacobian frozen at the step's start and each linear system W dz = -G,
    W = I - d h J, by restarted GMRES on fvm_jvp (matrix-free: no Jacobian is
    formed, so `linsolve` has no counterpart).  The local error estimate
    k h (f_n / gamma - f_gamma / (gamma (1 - gamma)) + f_{n+1} / (1 - gamma)),
    k = (3 gamma^2 - 4 gamma + 2) / (6 (2 - gamma)), is filtered through W^-1 for
    stiff problems (smooth_est), and the step follows the standard controller
    dt *= clamp(0.9 EEst^(-1/3), 0.2, 5)."""
    newton_kappa: float = 1e-2     # stop Newton when theta / (1 - theta) |dz| <= kappa (error norm)
    max_newton: int = 10
    krylov_rtol: float = 1e-3
    krylov_restart: int = 30
    smooth_est: bool = True
    device_loop: bool = True       # small affine problems: the whole integration in one launch (implicit.py)
    precond: str = "jacobi"        # right-precondition GMRES with diag(W) (FVMProblem), or "none"


@dataclass
class ODESolution:
    t: list = field(default_factory=list)
    u: list = field(default_factory=list)
    retcode: str = "Default"
    stats: dict = field(default_factory=dict)

    def __getitem__(self, k):
        return self.u[k]

    def __len__(self):
        return len(self.u)


def ctrl_log(x):
    """log(x) as the library's step controllers compute it (fvm_ctrl.hpp, the same
    operations in the same order): x = m 2^e, m in [sqrt(1/2), sqrt(2)),
    s = (m - 1) / (m + 1), log m = 2 (s + s^3 / 3 + ... + s^25 / 25)."""
    m, e = math.frexp(x)
    if m < 0.70710678118654752:
        m = m * 2.0
        e = e - 1
    s = (m - 1.0) / (m + 1.0)
    s2 = s * s
    p = 1.0 / 25.0
    for k in range(23, 0, -2):
        p = p * s2 + 1.0 / float(k)
    return 2.0 * s * p + float(e) * 0.69314718055994531


def ctrl_exp(y):
    """exp(y) as fvm_ctrl.hpp computes it: y = k ln 2 + r, Taylor to r^20 / 20!, times 2^k."""
    k = math.floor(y / 0.69314718055994531 + 0.5)
    r = y - k * 0.69314718055994531
    p = 1.0
    for i in range(20, 0, -1):
        p = 1.0 + p * r / float(i)
    return math.ldexp(p, int(k))


def ctrl_pow(x, p):
    """x^p of the step-size controllers: the library's device and host loops and this
    package's host loop use the same operations, so they pick the same steps bit for
    bit (libm's and the device library's pow differ in the last place)."""
    return ctrl_exp(p * ctrl_log(x))


class _Saver:
    def __init__(self, saveat, t0, tf):
        if saveat is None or (np.ndim(saveat) == 0 and float(saveat) == 0):
            self.times = None
        elif np.ndim(saveat) == 0:
            h = float(saveat)
            n = int(math.floor((tf - t0) / h + 1e-9))
            ts = [t0 + k * h for k in range(n + 1)]
            if ts[-1] < tf - 1e-12 * max(1.0, abs(tf)):
                ts.append(tf)
            self.times = ts
        else:
            self.times = sorted(float(v) for v in saveat)
        self.k = 0


def _hermite(out, u0, f0, u1, f1, dt, theta):
    # explicit products (the library's integrator forms the same weights)
    t2 = theta * theta
    t3 = t2 * theta
    h00 = 2 * t3 - 3 * t2 + 1
    h10 = t3 - 2 * t2 + theta
    h01 = -2 * t3 + 3 * t2
    h11 = t3 - t2
    _lib.lincomb(out, (h00, h10 * dt, h01, h11 * dt), (u0, f0, u1, f1))


def solve(prob, alg=None, *, saveat=None, abstol=None, reltol=None, dt=None, maxiters=10**7,
          save_start=True, callback=None, parallel=True, **kwargs):
    """solve(prob, alg; saveat, abstol, reltol, dt, ...).

    FVMProblem / FVMSystem: integrates on the device (abstol 1e-6, reltol 1e-3 unless
    given, OrdinaryDiffEq's defaults); returns ODESolution (t, u host arrays).
    SteadyFVMProblem: see solve_steady (its own defaults).  Templates forward to
    their own solver."""
    from .ensemble import EnsembleProblem
    from .templates import AbstractFVMTemplate, _SteadyTemplate

    if isinstance(prob, EnsembleProblem):  # solve(ensemble_prob, alg; trajectories, ...)
        if saveat is not None:
            kwargs["saveat"] = saveat
        for k, v in (("abstol", abstol), ("reltol", reltol)):
            if v is not None:
                kwargs[k] = v
        return prob.solve(alg, dt=dt, **kwargs)
    if isinstance(prob, _SteadyTemplate):  # a LinearProblem: only the linear solver's options apply
        bad = [k for k, v in (("saveat", saveat), ("dt", dt), ("callback", callback), ("abstol", abstol))
               if v is not None]
        if bad or not save_start:
            raise TypeError(f"solve: a steady template (a LinearProblem) takes no {', '.join(bad) or 'save_start'}")
        if reltol is not None:  # LinearSolve's reltol is the Krylov stopping tolerance
            kwargs.setdefault("rtol", reltol)
        if maxiters != 10**7:
            kwargs.setdefault("maxiter", maxiters)
        return prob.solve(alg, **kwargs)
    if isinstance(prob, SteadyFVMProblem):
        tol = {k: v for k, v in (("abstol", abstol), ("reltol", reltol)) if v is not None}
        return solve_steady(prob, alg, **tol, **kwargs)
    abstol = 1e-6 if abstol is None else abstol
    reltol = 1e-3 if reltol is None else reltol
    if isinstance(prob, AbstractFVMTemplate):  # every option goes on to the ODEProblem's solve
        return prob.solve(alg, saveat=saveat, abstol=abstol, reltol=reltol, dt=dt, maxiters=maxiters,
                          save_start=save_start, callback=callback, **kwargs)
    if isinstance(prob, FVMSystem):
        ctx = prob.device_context()
        cb = (lambda u, t: ctx.apply_dirichlet(u, t)) if prob.has_dirichlet_nodes() else None
        sol = integrate(ctx.rhs, prob.flat(prob.initial_condition), prob.initial_time,
                        prob.final_time, alg if alg is not None else Tsit5(), saveat=saveat,
                        abstol=abstol, reltol=reltol, dt=dt, maxiters=maxiters,
                        save_start=save_start, step_callback=cb, user_callback=callback,
                        jvp=ctx.jvp, op=(_lib.OP_SYSTEM, ctx.handle, cb is not None), fsal=fsal_ok(prob))
        sol.u = [prob.unflat(v) for v in sol.u]
        return sol
    if not isinstance(prob, FVMProblem):
        raise TypeError(f"solve: unsupported problem type {type(prob).__name__}")
    ctx = prob.device_context()
    if isinstance(alg, TRBDF2):
        from . import implicit

        if implicit.eligible(prob, alg, saveat, callback):
            return implicit.trbdf2_affine(prob, alg, saveat, abstol, reltol, dt, maxiters, save_start)
    rhs = lambda u, du, t: ctx.rhs(u, du, t)  # noqa: E731
    has_dir = prob.conditions.has_dirichlet_nodes()
    cb = (lambda u, t: ctx.apply_dirichlet(u, t)) if has_dir else None
    extra = {}
    if isinstance(alg, TRBDF2) and alg.precond == "jacobi":
        extra["jac_diag"] = jacobian_diagonal(prob)
    elif isinstance(alg, TRBDF2) and alg.precond != "none":
        raise ValueError(f"TRBDF2: unknown precond {alg.precond!r} (\"jacobi\" or \"none\")")
    return integrate(rhs, prob.initial_condition, prob.initial_time, prob.final_time,
                     alg if alg is not None else Tsit5(), saveat=saveat, abstol=abstol,
                     reltol=reltol, dt=dt, maxiters=maxiters, save_start=save_start,
                     step_callback=cb, user_callback=callback,
                     jvp=lambda u, v, out, t: ctx.jvp(u, v, out, t),
                     op=(_lib.OP_PROBLEM, ctx.handle, has_dir), fsal=fsal_ok(prob), **extra)


def jacobian_diagonal(prob):
    """diag(u, t, out): out = the diagonal of fvm_eqs!'s Jacobian at (u, t), from one
    fvm_jvp per colour of a distance-2 colouring of jacobian_sparsity's pattern
    (columns of one colour share no row, so J p_c holds each of their diagonal entries
    in its own row: out = sum_c p_c * J p_c).  The colouring and the probe vectors
    are kept on the problem."""
    from .equations import csr_pattern
    from .implicit import _colouring

    ctx = prob.device_context()
    probes = getattr(prob, "_diag_probes", None)
    if probes is None:
        n = prob.mesh.triangulation.num_points()
        rp, col = csr_pattern(prob.mesh.triangulation)
        colour = _colouring(rp, col, n)
        probes = [_lib.DeviceArray.from_host((colour == c).astype(np.float64)) for c in range(int(colour.max()) + 1)]
        prob._diag_probes = probes
    jv = _lib.DeviceArray(probes[0].n)

    def diag(u, t, out):
        out.fill_zero()
        for p in probes:
            ctx.jvp(u, p, jv, t)
            _lib.mul(jv, p, jv)
            _lib.lincomb(out, (1.0, 1.0), (out, jv))
    return diag


def fsal_ok(prob):
    """Whether Tsit5 may reuse its last stage as f(u) after an accepted step (first
    same as last) for this FVMProblem or FVMSystem: the Dirichlet update after each step can change
    u only where its values depend on t or u, or at the first step (the initial
    condition need not hold them); du is 0 at Dirichlet nodes, so a stage never moves
    them.  True without Dirichlet nodes or when every Dirichlet function is constant in
    t and u."""
    if isinstance(prob, FVMSystem):  # per-variable codes, condition forms over all variables
        if not prob.has_dirichlet_nodes():
            return True
        ph = prob.device_context().physics
        used = np.unique(ph["node_fidx"][ph["node_kind"] == _lib.NODE_DIRICHLET])
        return all(not np.any(np.asarray(ph[k])[used]) for k in ("fn_cu", "fn_ct", "fn_lam"))
    if not prob.conditions.has_dirichlet_nodes():
        return True
    fn = prob.device_context().physics["fn"]
    kind, fidx, _ = prob.conditions.flatten(prob.mesh.triangulation)
    used = np.unique(fidx[kind == _lib.NODE_DIRICHLET])
    coef = [np.asarray(fn[1]), np.asarray(fn[2])] + ([np.asarray(fn[5])] if len(fn) > 5 else [])
    return all(float(c[f]) == 0.0 for c in coef for f in used)


def integrate(rhs, u0, t0, tf, alg, *, saveat=None, save_start=True, **kwargs):
    """Integrate du/dt = rhs(u, t) on the device from t0 to tf (see _integrate).
    save_start=False drops the initial state from the saved ones, also when it is one
    of the saveat times (OrdinaryDiffEq's save_start)."""
    sol = _integrate(rhs, u0, t0, tf, alg, saveat=saveat, save_start=save_start, **kwargs)
    if not save_start and saveat is not None:
        keep = [k for k, tt in enumerate(sol.t) if tt > t0]
        sol.t = [sol.t[k] for k in keep]
        sol.u = [sol.u[k] for k in keep]
    return sol


def _integrate(rhs, u0, t0, tf, alg, *, saveat=None, abstol=1e-6, reltol=1e-3, dt=None,
               maxiters=10**7, save_start=True, step_callback=None, user_callback=None, jvp=None,
               op=None, fsal=False, jac_diag=None):
    """Integrate du/dt = rhs(u, t) on the device from t0 to tf.  Implicit methods
    need jvp(u, v, out, t) = J(u) v.

    fsal: step_callback cannot change u after the first accepted step (see
    fsal_ok), so Tsit5 takes f(u) after an accepted step from its last stage instead
    of evaluating it again, unless a user callback runs (it may change u).

    op = (kind, handle, dirichlet): the library operator that rhs evaluates (and,
    with dirichlet, whose Dirichlet update step_callback applies).  Explicit
    methods with saveat times and no user callback then run whole in the library
    (fvm_ode_solve: the same arithmetic, one call instead of one per stage)."""
    if (op is not None and user_callback is None and isinstance(alg, (Tsit5, Euler, SSPRK3))
            and saveat is not None):
        times = _Saver(saveat, t0, tf).times
        if times is not None and 8 * len(times) * len(u0) <= _DEVICE_SAVE_BYTES:
            return _integrate_device(op, u0, t0, tf, alg, times, abstol, reltol, dt, maxiters, fsal)
    if isinstance(alg, TRBDF2):
        if jvp is None:
            raise TypeError("TRBDF2 needs the Jacobian-vector product of the right-hand side")
        return _integrate_trbdf2(rhs, jvp, u0, t0, tf, alg, saveat=saveat, abstol=abstol,
                                 reltol=reltol, dt=dt, maxiters=maxiters, save_start=save_start,
                                 step_callback=step_callback, user_callback=user_callback, op=op,
                                 jac_diag=jac_diag)
    n = len(u0)
    u = _lib.DeviceArray.from_host(np.asarray(u0, np.float64))
    sol = ODESolution()
    saver = _Saver(saveat, t0, tf)
    nf = [0]

    def f(uu, out, tt):
        nf[0] += 1
        rhs(uu, out, tt)

    def save(tt, arr):
        sol.t.append(tt)
        sol.u.append(arr.to_host())

    if saver.times is None:
        if save_start:
            save(t0, u)
    t = t0
    naccept = nreject = 0
    if isinstance(alg, Tsit5):
        k = [_lib.DeviceArray(n) for _ in range(7)]
        utmp = _lib.DeviceArray(n)
        unew = _lib.DeviceArray(n)
        interp = _lib.DeviceArray(n)
        f(u, k[0], t)
        if dt is None:
            dt = _initial_dt(f, u, k[0], t, tf, abstol, reltol, n)
        qold = 1e-4
        it = 0
        if saver.times is not None:
            while saver.k < len(saver.times) and saver.times[saver.k] <= t + 1e-14 * max(1, abs(t)):
                save(saver.times[saver.k], u)
                saver.k += 1
        while t < tf and it < maxiters:
            it += 1
            dt = min(dt, tf - t)
            for s in range(1, 7):
                coefs = [1.0] + [dt * a for a in _TSIT5_A[s]]
                xs = [u] + k[:s]
                dst = unew if s == 6 else utmp
                _lib.lincomb(dst, coefs, xs)
                f(dst, k[s], t + _TSIT5_C[s] * dt)
            EEst = _lib.lincomb_err_norm([dt * b for b in _TSIT5_BTILDE], k, u, unew, abstol, reltol)
            if not np.isfinite(EEst):
                sol.retcode = "Unstable"
                break
            if EEst <= 1.0:
                # dense output for saveat inside (t, t + dt]
                tn = t + dt
                if saver.times is not None:
                    while saver.k < len(saver.times) and saver.times[saver.k] <= tn + 1e-12 * max(1, abs(tn)):
                        ts = saver.times[saver.k]
                        theta = min(1.0, max(0.0, (ts - t) / dt))
                        _hermite(interp, u, k[0], unew, k[6], dt, theta)
                        save(ts, interp)
                        saver.k += 1
                u, unew = unew, u
                t = tn
                naccept += 1
                if step_callback is not None:
                    step_callback(u, t)
                if user_callback is not None:
                    user_callback(u, t)
                if saver.times is None:
                    save(t, u)
                if fsal and user_callback is None and (step_callback is None or naccept >= 2):
                    k[0], k[6] = k[6], k[0]  # first same as last
                else:
                    f(u, k[0], t)
                q = ctrl_pow(max(EEst, 1e-16), 7 / 50) / ctrl_pow(qold, 2 / 25)
                q = min(5.0, max(0.1, q / 0.9))
                dt = dt / q
                qold = max(EEst, 1e-4)
            else:
                nreject += 1
                q = min(5.0, ctrl_pow(max(EEst, 1e-16), 7 / 50) / 0.9)
                dt = dt / q
        sol.retcode = "Success" if t >= tf else (sol.retcode if sol.retcode != "Default" else "MaxIters")
    else:
        if dt is None and getattr(alg, "dt", None) is None:
            raise ValueError("fixed-step integrators need dt")
        h = float(dt if dt is not None else alg.dt)
        nsteps = int(math.ceil((tf - t0) / h - 1e-9))
        k1 = _lib.DeviceArray(n)
        u1 = _lib.DeviceArray(n)
        u2 = _lib.DeviceArray(n)
        if saver.times is not None:
            while saver.k < len(saver.times) and saver.times[saver.k] <= t + 1e-14:
                save(saver.times[saver.k], u)
                saver.k += 1
        for step in range(nsteps):
            hh = min(h, tf - t)
            if isinstance(alg, Euler):
                f(u, k1, t)
                _lib.lincomb(u1, (1.0, hh), (u, k1))
            elif isinstance(alg, SSPRK3):
                f(u, k1, t)
                _lib.lincomb(u1, (1.0, hh), (u, k1))
                f(u1, k1, t + hh)
                _lib.lincomb(u2, (0.75, 0.25, 0.25 * hh), (u, u1, k1))
                f(u2, k1, t + 0.5 * hh)
                _lib.lincomb(u1, (1 / 3, 2 / 3, 2 / 3 * hh), (u, u2, k1))
            else:
                raise TypeError(f"unsupported algorithm {type(alg).__name__}")
            tn = t + hh
            if saver.times is not None:
                while saver.k < len(saver.times) and saver.times[saver.k] <= tn + 1e-12:
                    ts = saver.times[saver.k]
                    theta = (ts - t) / hh
                    _lib.lincomb(u2, (1 - theta, theta), (u, u1))
                    save(ts, u2)
                    saver.k += 1
            u, u1 = u1, u
            t = tn
            naccept += 1
            if step_callback is not None:
                step_callback(u, t)
            if user_callback is not None:
                user_callback(u, t)
            if saver.times is None:
                save(t, u)
        sol.retcode = "Success"
    sol.stats = dict(nf=nf[0], naccept=naccept, nreject=nreject)
    sol.final_device = u
    return sol


_TRB_G = 2.0 - math.sqrt(2.0)
_TRB_D = _TRB_G / 2                                    # = (1 - gamma) / (2 - gamma)
_TRB_A = 1.0 / (_TRB_G * (2.0 - _TRB_G))               # BDF2 weight of u_gamma
_TRB_B = -(1.0 - _TRB_G) ** 2 / (_TRB_G * (2.0 - _TRB_G))  # and of u_n
_TRB_K = (3 * _TRB_G ** 2 - 4 * _TRB_G + 2) / (6 * (2.0 - _TRB_G))


def _integrate_trbdf2(rhs, jvp, u0, t0, tf, alg, *, saveat, abstol, reltol, dt, maxiters,
                      save_start, step_callback, user_callback, op=None, jac_diag=None):
    """jac_diag(u, t, out): the diagonal of J(u, t) on the device; with it, GMRES is
    right-preconditioned by diag(W)^-1, W = I - d h J frozen at the step's start (one
    diagonal per step, its inverse per step size)."""
    from .linsolve import gmres_device, library_arnoldi

    n = len(u0)
    u = _lib.DeviceArray.from_host(np.asarray(u0, np.float64))
    sol = ODESolution()
    saver = _Saver(saveat, t0, tf)
    stats = dict(nf=0, njvp=0, naccept=0, nreject=0, nnewton=0, nkrylov=0, nnewton_fail=0)
    V = {k: _lib.DeviceArray(n) for k in ("fn", "z", "fg", "u1", "f1", "r", "G", "dz", "est", "tmp", "interp")}

    def f(uu, out, tt):
        stats["nf"] += 1
        rhs(uu, out, tt)

    def save(tt, arr):
        sol.t.append(tt)
        sol.u.append(arr.to_host())

    def W_apply(ubase, tbase, h):
        def apply(v, out):
            stats["njvp"] += 1
            jvp(ubase, v, out, tbase)
            _lib.lincomb(out, (1.0, -_TRB_D * h), (v, out))
        return apply

    def W_arnoldi(ubase, tbase, h):
        """The Arnoldi step for W = I - d h J in the library (op given), else None."""
        if op is None:
            return None
        step = library_arnoldi(op[0], op[1], ubase, tbase, 1.0, -_TRB_D * h)

        def counted(V, j, w, *pre):
            stats["njvp"] += 1
            return step(V, j, w, *pre)
        return counted

    # Jacobi preconditioner: diag(J) at the step's start, downloaded once per step;
    # 1 / diag(W) for each step size tried from it
    pre = dict(t=None, d=None, h=None)
    S = _lib.DeviceArray(n) if jac_diag is not None else None

    def precond(ubase, tbase, h):
        if jac_diag is None:
            return None
        if pre["t"] != tbase or pre["d"] is None:
            jac_diag(ubase, tbase, V["tmp"])
            pre["d"], pre["t"], pre["h"] = V["tmp"].to_host(), tbase, None
        if pre["h"] != h:
            w = 1.0 - _TRB_D * h * pre["d"]
            s = np.where(np.abs(w) > 1e-12, 1.0 / np.where(w == 0.0, 1.0, w), 1.0)
            S.upload(s)
            pre["h"] = h
        return S

    nl = dict(eta=1.0)  # Newton's contraction estimate theta / (1 - theta), kept between solves

    def newton(z, r, ts, ubase, tbase, h, fz):
        """z <- solution of z - d h f(z, ts) = r (z holds the predictor); fz <- f(z).
        Stops when eta |dz| <= kappa (Hairer & Wanner, Solving ODEs II, IV.8), eta
        starting from the previous solve's; returns False when Newton diverges or
        does not converge in max_newton iterations."""
        A = W_apply(ubase, tbase, h)
        eta = ctrl_pow(max(nl["eta"], 1e-16), 0.8)
        prev = None
        for _ in range(alg.max_newton):
            f(z, fz, ts)
            _lib.lincomb(V["G"], (-1.0, _TRB_D * h, 1.0), (z, fz, r))  # -G = r + d h f(z) - z
            stats["nkrylov"] += gmres_device(A, V["G"], V["dz"], rtol=alg.krylov_rtol,
                                             restart=alg.krylov_restart, maxiter=20 * alg.krylov_restart,
                                             arnoldi=W_arnoldi(ubase, tbase, h),
                                             precond=precond(ubase, tbase, h))
            _lib.lincomb(z, (1.0, 1.0), (z, V["dz"]))
            stats["nnewton"] += 1
            ndz = _lib.err_norm(V["dz"], ubase, z, abstol, reltol)
            if not math.isfinite(ndz):
                return False
            if prev is not None:
                theta = ndz / prev
                if theta >= 1.0:
                    return False
                eta = theta / (1.0 - theta)
            if eta * ndz <= alg.newton_kappa or ndz == 0.0:
                break
            prev = ndz
        else:
            return False
        nl["eta"] = eta
        f(z, fz, ts)
        return True

    t = t0
    f(u, V["fn"], t)
    if saver.times is None:
        if save_start:
            save(t0, u)
    else:
        while saver.k < len(saver.times) and saver.times[saver.k] <= t + 1e-14 * max(1, abs(t)):
            save(saver.times[saver.k], u)
            saver.k += 1
    if dt is None:
        dt = _initial_dt(f, u, V["fn"], t, tf, abstol, reltol, n, exponent=1 / 3)
    it = 0
    while t < tf and it < maxiters:
        it += 1
        h = min(dt, tf - t)
        fn, z, fg, u1, f1 = V["fn"], V["z"], V["fg"], V["u1"], V["f1"]
        # TR stage: z - d h f(z) = u + d h f_n, predictor u + gamma h f_n
        _lib.lincomb(V["r"], (1.0, _TRB_D * h), (u, fn))
        _lib.lincomb(z, (1.0, _TRB_G * h), (u, fn))
        ok = newton(z, V["r"], t + _TRB_G * h, u, t, h, fg)
        if ok:
            # BDF2 stage: u1 - d h f(u1) = a z + b u, predictor z + (1 - gamma) h f_gamma
            _lib.lincomb(V["r"], (_TRB_A, _TRB_B), (z, u))
            _lib.lincomb(u1, (1.0, (1.0 - _TRB_G) * h), (z, fg))
            ok = newton(u1, V["r"], t + h, u, t, h, f1)
        if not ok:
            stats["nnewton_fail"] += 1
            stats["nreject"] += 1
            dt = h / 4
            if dt < 1e-14 * max(1.0, abs(t)):
                sol.retcode = "ConvergenceFailure"
                break
            continue
        est = V["est"]
        _lib.lincomb(est, (_TRB_K * h / _TRB_G, -_TRB_K * h / (_TRB_G * (1 - _TRB_G)),
                           _TRB_K * h / (1 - _TRB_G)), (fn, fg, f1))
        if alg.smooth_est:  # W^-1 est (Hosea & Shampine): keeps stiff components from inflating it
            stats["nkrylov"] += gmres_device(W_apply(u, t, h), est, V["tmp"], rtol=alg.krylov_rtol,
                                             restart=alg.krylov_restart, maxiter=20 * alg.krylov_restart,
                                             arnoldi=W_arnoldi(u, t, h), precond=precond(u, t, h))
            est = V["tmp"]
        EEst = _lib.err_norm(est, u, u1, abstol, reltol)
        if not math.isfinite(EEst):
            sol.retcode = "Unstable"
            break
        if EEst <= 1.0:
            tn = t + h
            if saver.times is not None:
                while saver.k < len(saver.times) and saver.times[saver.k] <= tn + 1e-12 * max(1, abs(tn)):
                    ts = saver.times[saver.k]
                    theta = min(1.0, max(0.0, (ts - t) / h))
                    _hermite(V["interp"], u, fn, u1, f1, h, theta)
                    save(ts, V["interp"])
                    saver.k += 1
            u.copy_from(u1)
            t = tn
            stats["naccept"] += 1
            if step_callback is not None:
                step_callback(u, t)
            if user_callback is not None:
                user_callback(u, t)
            if saver.times is None:
                save(t, u)
            if step_callback is not None or user_callback is not None:
                f(u, fn, t)  # the callback may have changed u
            else:
                fn.copy_from(f1)
        else:
            stats["nreject"] += 1
        q = 0.9 * ctrl_pow(max(EEst, 1e-10), -1.0 / 3.0)
        dt = h * min(5.0, max(0.2, q))
    if sol.retcode == "Default":
        sol.retcode = "Success" if t >= tf else "MaxIters"
    sol.stats = stats
    sol.final_device = u
    return sol


_DEVICE_SAVE_BYTES = 8 << 30  # saved states held on the device by fvm_ode_solve


def _integrate_device(op, u0, t0, tf, alg, times, abstol, reltol, dt, maxiters, fsal=False):
    """The explicit loops of integrate() run by fvm_ode_solve; the saved states are
    written to one device buffer and copied back once."""
    kind, handle, dirichlet = op
    n = len(u0)
    o = _lib.OdeOpts()
    if isinstance(alg, Tsit5):
        o.method = _lib.ODE_TSIT5
        o.dt = float(dt) if dt is not None else 0.0
    else:
        h = dt if dt is not None else alg.dt
        if h is None:
            raise ValueError("fixed-step integrators need dt")
        o.method = _lib.ODE_EULER if isinstance(alg, Euler) else _lib.ODE_SSPRK3
        o.dt = float(h)
    u = _lib.DeviceArray.from_host(np.asarray(u0, np.float64))
    ts = np.ascontiguousarray(times, np.float64)
    saves = _lib.DeviceArray(max(len(ts), 1) * n)
    o.op_kind, o.op, o.dirichlet = int(kind), handle, int(bool(dirichlet))
    o.fsal = int(bool(fsal))
    o.n, o.t0, o.tf = n, float(t0), float(tf)
    o.abstol, o.reltol, o.maxiters = float(abstol), float(reltol), int(maxiters)
    o.n_save, o.save_t, o.d_saves = len(ts), _lib.ptr(ts), saves.ptr
    st = _lib.OdeStats()
    _lib.check(_lib.hip_lib().fvm_ode_solve(ctypes.byref(o), u.ptr, ctypes.byref(st), None), "fvm_ode_solve")
    host = saves.to_host()
    sol = ODESolution()
    sol.t = [float(v) for v in ts[:st.nsaved]]
    sol.u = [host[k * n:(k + 1) * n].copy() for k in range(st.nsaved)]
    sol.retcode = _lib.ODE_RETCODES[st.retcode]
    sol.stats = dict(nf=st.nf, naccept=st.naccept, nreject=st.nreject)
    sol.final_device = u
    return sol


def _initial_dt(f, u0, f0, t0, tf, abstol, reltol, n, exponent=1 / 5):
    """Hairer, Norsett & Wanner (1993), II.4: starting step size, (0.01 / m)^exponent
    (1/5 for Tsit5 here, 1/3 for TR-BDF2)."""
    tmp = _lib.DeviceArray(n)
    f1 = _lib.DeviceArray(n)
    d0 = _lib.err_norm(u0, u0, u0, abstol, reltol)
    d1 = _lib.err_norm(f0, u0, u0, abstol, reltol)
    dt0 = 1e-6 if (d0 < 1e-5 or d1 < 1e-5) else 0.01 * d0 / d1
    dt0 = min(dt0, tf - t0)
    _lib.lincomb(tmp, (1.0, dt0), (u0, f0))
    f(tmp, f1, t0 + dt0)
    _lib.lincomb(tmp, (1.0 / dt0, -1.0 / dt0), (f1, f0))
    d2 = _lib.err_norm(tmp, u0, u0, abstol, reltol)
    m = max(d1, d2)
    dt1 = max(1e-6, dt0 * 1e-3) if m <= 1e-15 else (0.01 / m) ** exponent
    return min(100 * dt0, dt1, tf - t0)


def _steady_affine(p, u, abstol, reltol, r0):
    """F(u) = J u + c = 0 for an FVMProblem affine in u, as one linear system on the
    device: J from coloured fvm_jvp probes (implicit.affine_operator, kept on the
    problem), A = J with the rows whose F is identically zero (Dirichlet nodes, points
    that are not vertices) made x_i = u_i, b = -c, solved by the template solvers'
    mixed-precision CG (fvm_template_create_csr, fvm_template_solve) to
    |b - A x| = |F(x)| <= abstol + reltol |F(u0)|, or, when CG stops short (it needs
    -diag(V) A close to symmetric positive definite, which a reaction term can spoil),
    by their BiCGStab.  Returns (x, stats, method), or None when neither converged (the
    caller then runs Newton-GMRES from u)."""
    from .implicit import affine_operator
    from .templates import BiCGStab, DeviceTemplate, MixedCG

    A = affine_operator(p)
    rp, col, val = A.csr
    if getattr(A, "_fixed", None) is None:
        n = len(rp) - 1
        c = A.d_c.to_host()
        rows = np.repeat(np.arange(n), np.diff(rp))
        fixed = np.bincount(rows, weights=np.abs(val), minlength=n) == 0
        fixed &= c == 0.0
        fixed[A.dir_nodes] = True
        A._fixed, A._minus_c = fixed, -c
    fixed = A._fixed
    b = np.where(fixed, u.to_host(), A._minus_c)
    key = (id(A), b.tobytes())
    cached = getattr(p, "_steady_linear", None)
    if cached is None or cached[0] != key:
        cached = (key, DeviceTemplate.from_csr(p.mesh, rp, col, val, b, fixed))
        p._steady_linear = cached
    T = cached[1]
    bn = float(np.linalg.norm(b))
    rtol = max((abstol + reltol * r0) / bn, 1e-15) if bn > 0 else 1e-15
    for alg in (MixedCG(), BiCGStab()):
        x = _lib.DeviceArray(len(b))
        _lib.lincomb(x, (1.0,), (u,))
        x, st = T.solve(alg, rtol=min(rtol, 1e-2), x=x, maxiter=20 * len(b) + 1000)
        if st["converged"]:
            return x, st, f"{type(alg).__name__} on the extracted Jacobian"
    return None


def solve_steady(prob, alg=None, *, abstol=1e-8, reltol=1e-8, maxiters=200, precond="jacobi", linear="auto",
                 **kwargs):
    """SteadyFVMProblem: Newton's method on fvm_eqs!(du, u) = 0 with the device
    Jacobian-vector product (fvm_jvp, fvm_system_jvp) inside restarted GMRES.  Dirichlet nodes are
    set by the Dirichlet functions first and stay fixed (their du is 0).  Stops when
    |F| <= abstol + reltol |F0| and the last correction is below sqrt(reltol) (1 + |u|).
    precond="jacobi" (FVMProblem): GMRES right-preconditioned with diag(J)^-1, the
    diagonal from coloured fvm_jvp probes (jacobian_diagonal), taken once for a problem
    affine in u and at every Newton iterate otherwise (rows with a zero diagonal, the
    Dirichlet nodes', scaled by 1); "none": plain GMRES.  linear="auto": an FVMProblem
    affine in u is first solved as one linear system with the template solvers
    (_steady_affine), Newton-GMRES only finishing what that leaves; "newton": Newton-GMRES
    throughout."""
    from .implicit import is_affine
    from .linsolve import gmres_device, library_arnoldi

    if precond not in ("jacobi", "none"):
        raise ValueError(f"solve: unknown precond {precond!r} (jacobi or none)")
    if linear not in ("auto", "newton"):
        raise ValueError(f"solve: unknown linear {linear!r} (auto or newton)")
    p = prob.problem
    ctx = p.device_context()
    sys = isinstance(p, FVMSystem)
    u0 = p.flat(p.initial_condition) if sys else p.initial_condition
    n = len(u0)
    u = _lib.DeviceArray.from_host(u0)
    if p.has_dirichlet_nodes():
        ctx.apply_dirichlet(u, p.final_time if np.isfinite(p.final_time) else 0.0)
    t = 0.0
    F = _lib.DeviceArray(n)
    du = _lib.DeviceArray(n)
    sol = ODESolution()
    ctx.rhs(u, F, t)
    r0 = math.sqrt(_lib.dot(F, F))
    it = 0
    rn = r0
    step = math.inf
    nkrylov = 0
    affine = not sys and is_affine(p)
    lin = None
    if affine and linear == "auto" and rn > abstol:
        lin = _steady_affine(p, u, abstol, reltol, r0)
        if lin is not None:
            u = lin[0]
            ctx.rhs(u, F, t)
            rn = math.sqrt(_lib.dot(F, F))
    jdiag = S = None  # the Jacobi preconditioner, made when Newton-GMRES first runs

    def converged():
        # the residual has dropped, and the last Newton correction is small next to u
        # (GMRES solves each correction to 1e-3 only; a linear solve to the tolerance
        # needs no correction)
        if rn > abstol + reltol * r0:
            return False
        return it == 0 or step <= math.sqrt(reltol) * (1.0 + math.sqrt(_lib.dot(u, u)))

    while it < maxiters:
        if converged():
            break
        it += 1
        if it == 1 and precond == "jacobi" and not sys:
            jdiag, S = jacobian_diagonal(p), _lib.DeviceArray(n)
        if jdiag is not None and (it == 1 or not affine):
            jdiag(u, t, du)
            d = du.to_host()
            S.upload(np.where(d != 0.0, 1.0 / np.where(d == 0.0, 1.0, d), 1.0))
        jv = lambda v, out: ctx.jvp(u, v, out, t)  # noqa: E731
        nkrylov += gmres_device(jv, F, du, rtol=1e-3, restart=50, maxiter=2000, precond=S,
                                arnoldi=library_arnoldi(_lib.OP_SYSTEM if sys else _lib.OP_PROBLEM, ctx.handle, u, t,
                                                        0.0, 1.0))
        _lib.lincomb(u, (1.0, -1.0), (u, du))
        step = math.sqrt(_lib.dot(du, du))
        ctx.rhs(u, F, t)
        rn = math.sqrt(_lib.dot(F, F))
    sol.t = [math.inf]
    sol.u = [p.unflat(u.to_host()) if sys else u.to_host()]
    sol.retcode = "Success" if converged() else "MaxIters"
    sol.stats = dict(newton_iters=it, residual=rn, initial_residual=r0, last_step=step, nkrylov=nkrylov,
                     precond=precond if jdiag is not None else "none")
    if lin is not None:
        sol.stats["linear_solve"] = dict(method=lin[2], iterations=lin[1]["iterations"], outer=lin[1]["outer"])
    return sol
