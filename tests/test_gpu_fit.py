"""Discrete-adjoint gradients of a trajectory loss (adjoint.py) and the nodal log-D
fit: the gradient against central differences of the loss (step 1e-6, bar 1e-6),
checkpointing that recomputes states giving the same gradient bit for bit, and
gradient descent recovering a loss drop on a synthetic target."""
import numpy as np
import pytest

import finitevolumemethod as FVM
from finitevolumemethod import adjoint as AD
from finitevolumemethod import ensemble as EN

pytestmark = pytest.mark.gpu


def setup(n=18, seed=3, consistent=False):
    """A nodal log-D problem with homogeneous Dirichlet conditions; with consistent=True
    the initial condition already holds them on the boundary."""
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, n, n, jitter=0.25, seed=seed)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    rng = np.random.default_rng(seed)
    N = tri.num_points()
    true = np.log(0.1) + 0.3 * rng.standard_normal(N)
    u0 = np.sin(np.pi * tri.x) * np.sin(np.pi * tri.y) + 0.2 * rng.random(N)
    if consistent:
        u0[(tri.x == 0.0) | (tri.x == 1.0) | (tri.y == 0.0) | (tri.y == 1.0)] = 0.0
    prob = FVM.FVMProblem(mesh, BCs, diffusion_function=FVM.NodalDiffusion(true),
                          initial_condition=u0, final_time=0.1)
    proxy = FVM.FVMProblem(mesh, BCs, diffusion_function=lambda x, y, t, u, p: float(np.exp(true.max())) + 0.0 * u,
                           initial_condition=u0, final_time=0.1)
    dt = 0.5 * EN.explicit_dt(proxy)
    return tri, prob, true, dt


def test_gradient_matches_central_differences(gpu):
    tri, prob, true, dt = setup()
    N = tri.num_points()
    fit = AD.DiffusivityFit(prob, dt, 30)
    target = fit.forward(true)
    rng = np.random.default_rng(9)
    logd = true + 0.2 * rng.standard_normal(N)
    loss, g, gu0 = fit.loss_and_gradient(logd, target)
    assert loss > 0
    for _ in range(3):
        z = rng.standard_normal(N)
        h = 1e-6
        lp = 0.5 * np.sum((fit.forward(logd + h * z) - target) ** 2)
        lm = 0.5 * np.sum((fit.forward(logd - h * z) - target) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(np.dot(g, z) - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(g) * np.linalg.norm(z))
    # the u0 gradient
    z = rng.standard_normal(N)
    h = 1e-6
    u0 = prob.initial_condition
    lp = 0.5 * np.sum((fit.forward(logd, u0 + h * z) - target) ** 2)
    lm = 0.5 * np.sum((fit.forward(logd, u0 - h * z) - target) ** 2)
    fd = (lp - lm) / (2 * h)
    assert abs(np.dot(gu0, z) - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(gu0) * np.linalg.norm(z))


def test_checkpointing_gives_the_same_gradient(gpu):
    tri, prob, true, dt = setup(n=12)
    logd = true + 0.1
    target = AD.DiffusivityFit(prob, dt, 17).forward(true)
    a = AD.DiffusivityFit(prob, dt, 17).loss_and_gradient(logd, target)
    b = AD.DiffusivityFit(prob, dt, 17, checkpoint=5).loss_and_gradient(logd, target)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_gradient_descent_reduces_the_loss(gpu):
    tri, prob, true, dt = setup(n=16)
    fit = AD.DiffusivityFit(prob, dt, 25)
    target = fit.forward(true)
    logd, hist = AD.fit_log_diffusivity(prob, target, np.full(len(true), np.log(0.1)), dt, 25, iters=50)
    assert len(hist) == 50
    assert all(b <= a for a, b in zip(hist, hist[1:]))
    assert hist[-1] <= 0.05 * hist[0]


def _fd_check(fit, logd, target, u0, rng, ndir=3, h=1e-6):
    N = len(logd)
    loss, g, gu0 = fit.loss_and_gradient(logd, target, u0)
    for _ in range(ndir):
        z = rng.standard_normal(N)
        lp = 0.5 * np.sum((fit.forward(logd + h * z, u0) - target) ** 2)
        lm = 0.5 * np.sum((fit.forward(logd - h * z, u0) - target) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(np.dot(g, z) - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(g) * np.linalg.norm(z)), (
            np.dot(g, z), fd)
    z = rng.standard_normal(N)
    lp = 0.5 * np.sum((fit.forward(logd, u0 + h * z) - target) ** 2)
    lm = 0.5 * np.sum((fit.forward(logd, u0 - h * z) - target) ** 2)
    fd = (lp - lm) / (2 * h)
    assert abs(np.dot(gu0, z) - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(gu0) * np.linalg.norm(z))
    return loss, g


def test_rkc_gradient_matches_central_differences(gpu):
    """The RKC discrete adjoint: steps 20x the forward-Euler limit, stages from the
    Gershgorin bound; gradients in log D and u0 against central differences."""
    tri, prob, true, dt = setup()
    fit = AD.DiffusivityFit(prob, 20 * dt, 4, method="rkc")
    assert fit.stages >= 6
    target = fit.forward(true)
    rng = np.random.default_rng(19)
    logd = true + 0.2 * rng.standard_normal(len(true))
    _fd_check(fit, logd, target, prob.initial_condition, rng)
    # checkpointing every second step recomputes the states: the same gradient bit for bit
    a = fit.loss_and_gradient(logd, target)
    b = AD.DiffusivityFit(prob, 20 * dt, 4, checkpoint=2, method="rkc").loss_and_gradient(logd, target)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_rkc_reduces_to_the_forward_state_of_a_tight_euler_run(gpu):
    """RKC is a consistent second-order integrator: 5 RKC steps match 2000 forward-Euler
    steps over the same horizon to the Euler run's own first-order error.  The initial
    condition holds the Dirichlet values, so the callback after each step changes
    nothing and both runs integrate the same ODE."""
    tri, prob, true, dt = setup(n=14, consistent=True)
    T = 40 * dt
    rk = AD.DiffusivityFit(prob, T / 5, 5, method="rkc").forward(true)
    e1 = AD.DiffusivityFit(prob, T / 1000, 1000).forward(true)
    e2 = AD.DiffusivityFit(prob, T / 2000, 2000).forward(true)
    ref = 2 * e2 - e1  # Richardson extrapolation of Euler
    assert np.max(np.abs(e2 - ref)) > 0
    assert np.max(np.abs(rk - ref)) <= 2e-2 * np.max(np.abs(ref - prob.initial_condition))


def test_rkc_gradient_on_the_differentiation_workload(gpu):
    """The 1000 x 1000 per-node log-D workload (u0 ~ U[0,1), log D ~ N(log 0.1, 0.3^2)),
    two RKC steps of 2e-4: gradient in log D against central differences (step 1e-6)."""
    from finitevolumemethod import workloads as WL

    tri, prob, u, logd, w = WL.differentiation_problem(1000)
    fit = AD.DiffusivityFit(prob, 2e-4, 2, method="rkc")
    rng = np.random.default_rng(23)
    target = fit.forward(logd + 0.1 * rng.standard_normal(len(logd)))
    print(f"RKC stages {fit.stages}")
    _fd_check(fit, logd, target, u, rng, ndir=2)
