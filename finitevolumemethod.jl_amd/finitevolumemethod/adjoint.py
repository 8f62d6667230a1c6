"""Gradients of a trajectory loss with respect to u0 and a per-node log-diffusivity
field, by the discrete adjoint of explicit steps built on fvm_vjp.

The reference differentiates forward only (ForwardDiff through DiffCache,
main_equations.jl:56); fitting a diffusivity field with it costs one tangent
solve per parameter.  Here one reverse sweep gives the whole gradient.

Forward Euler:

  forward   u_{k+1} = C(u_k + dt f(u_k, logD), t_{k+1})     (C: the Dirichlet callback)
  loss      L = 1/2 |u_K - target|^2
  reverse   lam_K = u_K - target
            lam~  = lam_{k+1} with its Dirichlet entries zeroed (C overwrites them)
            lam_k = lam~ + dt J_u(u_k)^T lam~,   g += dt J_logD(u_k)^T lam~

Runge-Kutta-Chebyshev (RKC, second order, s stages, damping eps; van der Houwen &
Sommeijer 1980, Sommeijer, Shampine & Verwer 1998), whose stability interval
[-beta, 0], beta ~ 0.653 s^2 at the published eps = 2/13 and ~ 0.347 s^2 at eps = 10,
makes stiff diffusion affordable explicitly (s grows with the square root of
dt rho(J)).  The fit uses eps = 10: every stiff mode is then damped to at most 0.24
per step (0.95 at eps = 2/13), so a rough initial state decays as it should:

  Y_0 = u_k,  Y_1 = Y_0 + mu~_1 dt F_0
  Y_j = (1 - mu_j - nu_j) Y_0 + mu_j Y_{j-1} + nu_j Y_{j-2} + mu~_j dt F_{j-1} + gamma~_j dt F_0
  u_{k+1} = C(Y_s),   F_j = f(Y_j)

  reverse, with lam_s = lam~ and the stage states of the step recomputed from u_k:
  for j = s..2:  lam_{j-1} += mu_j lam_j + mu~_j dt J_u(Y_{j-1})^T lam_j
                 lam_{j-2} += nu_j lam_j  (j > 2; nu_2 lam_2 goes to lam_0)
                 g += mu~_j dt J_logD(Y_{j-1})^T lam_j
  lam_0 = lam_1 + nu_2 lam_2 + sum_{j>=2} (1 - mu_j - nu_j) lam_j
  phi_0 = mu~_1 dt lam_1 + sum_{j>=2} gamma~_j dt lam_j
  lam_k = lam_0 + J_u(Y_0)^T phi_0,   g += J_logD(Y_0)^T phi_0

The forward states are kept on the device (checkpoint every `checkpoint` steps and
recompute the rest in the reverse sweep when memory matters).
"""
import math

import numpy as np

from . import _lib
from .forms import NodalDiffusion


def rkc_coefficients(s, eps=2.0 / 13.0):
    """The RKC(s) recurrence coefficients (Sommeijer, Shampine & Verwer 1998, eqs. 2.4-2.8):
    dict with mut1 and, for j = 2..s, mu[j], nu[j], mut[j], gt[j]; c[j] the stage times
    (fractions of dt); beta = (1 + w0) T_s''(w0) / T_s'(w0), the length of the real
    stability interval."""
    if s < 2:
        raise ValueError("RKC needs at least 2 stages")
    w0 = 1.0 + eps / s ** 2
    T, dT, ddT = [1.0, w0], [0.0, 1.0], [0.0, 0.0]
    for j in range(2, s + 1):
        T.append(2 * w0 * T[j - 1] - T[j - 2])
        dT.append(2 * T[j - 1] + 2 * w0 * dT[j - 1] - dT[j - 2])
        ddT.append(4 * dT[j - 1] + 2 * w0 * ddT[j - 1] - ddT[j - 2])
    w1 = dT[s] / ddT[s]
    b = [0.0] * (s + 1)
    for j in range(2, s + 1):
        b[j] = ddT[j] / dT[j] ** 2
    b[0] = b[1] = b[2]
    a = [1.0 - b[j] * T[j] for j in range(s + 1)]
    mu, nu, mut, gt = ([0.0] * (s + 1) for _ in range(4))
    for j in range(2, s + 1):
        mu[j] = 2 * w0 * b[j] / b[j - 1]
        nu[j] = -b[j] / b[j - 2]
        mut[j] = 2 * w1 * b[j] / b[j - 1]
        gt[j] = -a[j - 1] * mut[j]
    c = [0.0] * (s + 1)
    for j in range(2, s + 1):
        c[j] = w1 * ddT[j] / dT[j]
    c[1] = c[2] / dT[2]
    return dict(s=s, mut1=b[1] * w1, mu=mu, nu=nu, mut=mut, gt=gt, c=c,
                beta=(1 + w0) * ddT[s] / dT[s])


def rkc_stages(dt_rho, margin=1.25, eps=2.0 / 13.0):
    """The fewest RKC stages whose stability interval covers margin * dt * rho."""
    s = 2
    while rkc_coefficients(s, eps)["beta"] < margin * dt_rho:
        s += 1
    return s


def spectral_bound(prob, d_max):
    """Gershgorin bound of rho(J) for the problem's mesh and conditions with a constant
    diffusivity d_max (which bounds the nodal field's interpolant everywhere)."""
    from .ensemble import explicit_dt
    from .problem import FVMProblem

    unit = FVMProblem(prob.mesh, None, conditions=prob.conditions,
                      diffusion_function=lambda x, y, t, u, p: 1.0 + 0.0 * u,
                      initial_condition=prob.initial_condition, final_time=prob.final_time)
    return float(d_max) * 2.0 / explicit_dt(unit, safety=1.0)


class DiffusivityFit:
    """Loss and gradient of 1/2 |u(t_final) - target|^2 w.r.t. a nodal log D field for
    an FVMProblem whose diffusion_function is a NodalDiffusion, advanced by nsteps
    steps of size dt: forward Euler (method="euler"), or RKC with `stages` stages and
    damping `damping` (method="rkc"; by default the fewest stages that cover
    1.25 dt rho(J) for the problem's own log D field, rho from the Gershgorin bound)."""

    def __init__(self, prob, dt, nsteps, checkpoint=0, method="euler", stages=None, damping=10.0):
        if not isinstance(prob.diffusion_function, NodalDiffusion):
            raise TypeError("DiffusivityFit needs a problem with a NodalDiffusion diffusion_function")
        if method not in ("euler", "rkc"):
            raise ValueError("method must be 'euler' or 'rkc'")
        self.method = method
        self.prob = prob
        self.ctx = prob.device_context()
        self.n = len(prob.initial_condition)
        self.dt = float(dt)
        self.nsteps = int(nsteps)
        self.checkpoint = int(checkpoint) if checkpoint else 1
        kind, fidx, _ = prob.conditions.flatten(prob.mesh.triangulation)
        cu = np.asarray(self.ctx.physics["fn"][1])
        if np.any(cu[fidx[kind == _lib.NODE_DIRICHLET]] != 0.0):
            raise ValueError("DiffusivityFit needs Dirichlet values independent of u")
        self.dir_mask = _lib.DeviceArray.from_host((kind != _lib.NODE_DIRICHLET).astype(np.float64))
        self.has_dir = bool(np.any(kind == _lib.NODE_DIRICHLET))
        self.t0 = float(prob.initial_time)
        n = self.n
        self.d_logd = _lib.DeviceArray(n)
        self.f = _lib.DeviceArray(n)
        self.f2 = _lib.DeviceArray(n)
        self.ubar = _lib.DeviceArray(n)
        self.lbar = _lib.DeviceArray(n)
        self.lam = _lib.DeviceArray(n)
        self.grad = _lib.DeviceArray(n)
        ncp = math.ceil(self.nsteps / self.checkpoint) + 1
        self.states = [_lib.DeviceArray(n) for _ in range(ncp)]
        self.work = [_lib.DeviceArray(n) for _ in range(self.checkpoint)] if self.checkpoint > 1 else []
        if method == "rkc":
            logd0 = np.asarray(prob.diffusion_function.log_d, np.float64)
            self.rho_unit = spectral_bound(prob, 1.0)
            self.stages = int(stages) if stages else rkc_stages(
                self.dt * self.rho_unit * float(np.exp(logd0.max())), eps=damping)
            self.coef = rkc_coefficients(self.stages, damping)
            s = self.stages
            self.Y = [_lib.DeviceArray(n) for _ in range(s)]     # stage states of one step
            self.L = [_lib.DeviceArray(n) for _ in range(s + 1)]  # their adjoints
            self.F0 = _lib.DeviceArray(n)
            self.phi0 = _lib.DeviceArray(n)

    def rkc_stable(self, logd):
        """Whether the RKC stability interval covers dt rho(J) for this log D field."""
        return self.dt * self.rho_unit * float(np.exp(np.max(logd))) <= self.coef["beta"]

    def _step(self, u_in, u_out, k):
        if self.method == "rkc":
            return self._rkc_step(u_in, u_out, k)
        t = self.t0 + k * self.dt
        self.ctx.rhs(u_in, self.f, t)
        _lib.lincomb(u_out, (1.0, self.dt), (u_in, self.f))
        if self.has_dir:
            self.ctx.apply_dirichlet(u_out, t + self.dt)

    def _rkc_step(self, u_in, u_out, k, keep=False):
        """One RKC step u_in -> u_out (u_out must not alias u_in); keep=True leaves the
        stage states Y_0..Y_{s-1} in self.Y for the reverse sweep."""
        c, h, s = self.coef, self.dt, self.stages
        t = self.t0 + k * h
        Y = self.Y
        if keep:
            Y[0].copy_from(u_in)
            y0 = Y[0]
        else:
            y0 = u_in
        self.ctx.rhs(y0, self.F0, t)
        # without keep, the stages rotate through Y[0..2] (u_in stays Y_0)
        slot = (lambda j: Y[j]) if keep else (lambda j: Y[j % 3])
        prev2, prev1 = y0, (u_out if s == 1 else slot(1))
        _lib.lincomb(prev1, (1.0, c["mut1"] * h), (y0, self.F0))
        for j in range(2, s + 1):
            self.ctx.rhs(prev1, self.f, t + c["c"][j - 1] * h)
            dst = u_out if j == s else slot(j)
            _lib.lincomb(dst, (1.0 - c["mu"][j] - c["nu"][j], c["mu"][j], c["nu"][j], c["mut"][j] * h,
                               c["gt"][j] * h), (y0, prev1, prev2, self.f, self.F0))
            prev2, prev1 = prev1, dst
        if self.has_dir:
            self.ctx.apply_dirichlet(u_out, t + h)

    def _rkc_reverse(self, k):
        """lam (the adjoint of u_{k+1}, Dirichlet entries already zeroed) -> the adjoint
        of u_k, adding the step's log D gradient to grad.  The VJPs take lam_j itself
        and their results are scaled by mu~_j dt where they are added; the sums into
        lam_0 and phi_0, which every stage feeds, are formed once at the end over the
        kept stage adjoints."""
        c, h, s = self.coef, self.dt, self.stages
        t = self.t0 + k * h
        L, Y = self.L, self.Y
        self._rkc_step(self._state(k), self.f2, k, keep=True)  # stage states of step k
        L[s].copy_from(self.lam)
        for j in range(s):
            L[j].fill_zero()
        for j in range(s, 1, -1):
            m = c["mut"][j] * h
            self.ctx.vjp(Y[j - 1], L[j], self.ubar, self.lbar, t + c["c"][j - 1] * h)
            _lib.lincomb(L[j - 1], (1.0, c["mu"][j], m), (L[j - 1], L[j], self.ubar))
            if j > 2:
                _lib.lincomb(L[j - 2], (1.0, c["nu"][j]), (L[j - 2], L[j]))
            _lib.lincomb(self.grad, (1.0, m), (self.grad, self.lbar))
        # lam_0 = sum_{j>=2} (1 - mu_j - nu_j) lam_j + nu_2 lam_2 + lam_1 and
        # phi_0 = sum_{j>=2} gamma~_j dt lam_j + mu~_1 dt lam_1, 64 vectors per pass
        a0 = [1.0] + [1.0 - c["mu"][j] - c["nu"][j] + (c["nu"][j] if j == 2 else 0.0) for j in range(2, s + 1)]
        g0 = [c["mut1"] * h] + [c["gt"][j] * h for j in range(2, s + 1)]
        vecs = L[1:s + 1]
        self.phi0.fill_zero()
        for lo in range(0, s, 64):
            _lib.maxpy(L[0], a0[lo:lo + 64], vecs[lo:lo + 64])
            _lib.maxpy(self.phi0, g0[lo:lo + 64], vecs[lo:lo + 64])
        self.ctx.vjp(Y[0], self.phi0, self.ubar, self.lbar, t)
        _lib.lincomb(self.lam, (1.0, 1.0), (L[0], self.ubar))
        _lib.lincomb(self.grad, (1.0, 1.0), (self.grad, self.lbar))

    def forward(self, logd, u0=None):
        """Runs the steps; returns u(t_final) on the host."""
        if self.method == "rkc" and not self.rkc_stable(logd):
            raise ValueError(f"RKC with {self.stages} stages does not cover dt rho(J) for this log D "
                             "field; pass more stages")
        self.d_logd.upload(np.asarray(logd, np.float64))
        self.ctx.set_log_d(self.d_logd)
        cur = self.states[0]
        cur.upload(np.asarray(self.prob.initial_condition if u0 is None else u0, np.float64))
        u = _lib.DeviceArray(self.n)
        u.copy_from(cur)
        tmp = _lib.DeviceArray(self.n)
        for k in range(self.nsteps):
            self._step(u, tmp, k)
            u, tmp = tmp, u
            if (k + 1) % self.checkpoint == 0 and (k + 1) // self.checkpoint < len(self.states):
                self.states[(k + 1) // self.checkpoint].copy_from(u)
        self.final = u
        return u.to_host()

    def _state(self, k):
        """u_k on the device (recomputed from the previous checkpoint when needed)."""
        c, r = divmod(k, self.checkpoint)
        if r == 0:
            return self.states[c]
        u = self.work[0]
        u.copy_from(self.states[c])
        tmp = self.work[1] if len(self.work) > 1 else _lib.DeviceArray(self.n)
        for j in range(c * self.checkpoint, k):
            self._step(u, tmp, j)
            u.copy_from(tmp)
        return u

    def loss_and_gradient(self, logd, target, u0=None):
        """(L, dL/dlogD, dL/du0) as host arrays."""
        uK = self.forward(logd, u0)
        r = uK - np.asarray(target, np.float64)
        loss = 0.5 * float(np.dot(r, r))
        self.lam.upload(r)
        self.grad.fill_zero()
        for k in range(self.nsteps - 1, -1, -1):
            if self.has_dir:
                _lib.mul(self.lam, self.lam, self.dir_mask)
            if self.method == "rkc":
                self._rkc_reverse(k)
                continue
            uk = self._state(k)
            t = self.t0 + k * self.dt
            self.ctx.vjp(uk, self.lam, self.ubar, self.lbar, t)
            _lib.lincomb(self.lam, (1.0, self.dt), (self.lam, self.ubar))
            _lib.lincomb(self.grad, (1.0, self.dt), (self.grad, self.lbar))
        return loss, self.grad.to_host(), self.lam.to_host()


def fit_log_diffusivity(prob, target, logd0, dt, nsteps, iters=50, lr=None, checkpoint=0, method="euler",
                        stages=None, damping=10.0, fit=None):
    """Gradient descent on the nodal log D field (the diffusivity-field fit): returns
    (logD, loss history).  lr defaults to a backtracking step from 1.  A prepared
    DiffusivityFit may be passed as fit (its dt, steps and method are used)."""
    if fit is None:
        fit = DiffusivityFit(prob, dt, nsteps, checkpoint, method=method, stages=stages, damping=damping)
    logd = np.array(logd0, np.float64)
    hist = []
    step = 1.0 if lr is None else float(lr)
    for _ in range(iters):
        loss, g, _ = fit.loss_and_gradient(logd, target)
        hist.append(loss)
        if lr is not None:
            logd = logd - step * g
            continue
        gg = float(np.dot(g, g))
        if gg == 0.0:
            break
        while step > 1e-12:  # Armijo backtracking
            trial = logd - step * g
            if fit.method == "rkc" and not fit.rkc_stable(trial):
                step *= 0.5
                continue
            r = fit.forward(trial) - target
            if 0.5 * float(np.dot(r, r)) <= loss - 1e-4 * step * gg:
                logd = trial
                step *= 2.0
                break
            step *= 0.5
    return logd, hist
