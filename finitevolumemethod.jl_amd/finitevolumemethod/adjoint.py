"""Gradients of a trajectory loss with respect to u0 and a per-node log-diffusivity
field, by the discrete adjoint of forward-Euler steps built on fvm_vjp.

The reference differentiates forward only (ForwardDiff through DiffCache,
main_equations.jl:56); fitting a diffusivity field with it costs one tangent
solve per parameter.  Here one reverse sweep gives the whole gradient:

  forward   u_{k+1} = C(u_k + dt f(u_k, logD), t_{k+1})     (C: the Dirichlet callback)
  loss      L = 1/2 |u_K - target|^2
  reverse   lam_K = u_K - target
            lam~  = lam_{k+1} with its Dirichlet entries zeroed (C overwrites them)
            lam_k = lam~ + dt J_u(u_k)^T lam~,   g += dt J_logD(u_k)^T lam~

The forward states are kept on the device (checkpoint every `checkpoint` steps and
recompute the rest in the reverse sweep when memory matters).
"""
import math

import numpy as np

from . import _lib
from .forms import NodalDiffusion


class DiffusivityFit:
    """Loss and gradient of 1/2 |u(t_final) - target|^2 w.r.t. a nodal log D field for
    an FVMProblem whose diffusion_function is a NodalDiffusion, advanced by nsteps
    forward-Euler steps of size dt."""

    def __init__(self, prob, dt, nsteps, checkpoint=0):
        if not isinstance(prob.diffusion_function, NodalDiffusion):
            raise TypeError("DiffusivityFit needs a problem with a NodalDiffusion diffusion_function")
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
        self.ubar = _lib.DeviceArray(n)
        self.lbar = _lib.DeviceArray(n)
        self.lam = _lib.DeviceArray(n)
        self.grad = _lib.DeviceArray(n)
        ncp = math.ceil(self.nsteps / self.checkpoint) + 1
        self.states = [_lib.DeviceArray(n) for _ in range(ncp)]
        self.work = [_lib.DeviceArray(n) for _ in range(self.checkpoint)] if self.checkpoint > 1 else []

    def _step(self, u_in, u_out, k):
        t = self.t0 + k * self.dt
        self.ctx.rhs(u_in, self.f, t)
        _lib.lincomb(u_out, (1.0, self.dt), (u_in, self.f))
        if self.has_dir:
            self.ctx.apply_dirichlet(u_out, t + self.dt)

    def forward(self, logd, u0=None):
        """Runs the steps; returns u(t_final) on the host."""
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
            uk = self._state(k)
            t = self.t0 + k * self.dt
            self.ctx.vjp(uk, self.lam, self.ubar, self.lbar, t)
            _lib.lincomb(self.lam, (1.0, self.dt), (self.lam, self.ubar))
            _lib.lincomb(self.grad, (1.0, self.dt), (self.grad, self.lbar))
        return loss, self.grad.to_host(), self.lam.to_host()


def fit_log_diffusivity(prob, target, logd0, dt, nsteps, iters=50, lr=None, checkpoint=0):
    """Gradient descent on the nodal log D field (the diffusivity-field fit): returns
    (logD, loss history).  lr defaults to a backtracking step from 1."""
    fit = DiffusivityFit(prob, dt, nsteps, checkpoint)
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
            r = fit.forward(trial) - target
            if 0.5 * float(np.dot(r, r)) <= loss - 1e-4 * step * gg:
                logd = trial
                step *= 2.0
                break
            step *= 0.5
    return logd, hist
