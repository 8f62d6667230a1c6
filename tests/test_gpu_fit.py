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


def setup(n=18, seed=3):
    tri = FVM.jittered_rectangle(0.0, 1.0, 0.0, 1.0, n, n, jitter=0.25, seed=seed)
    mesh = FVM.FVMGeometry(tri)
    BCs = FVM.BoundaryConditions(mesh, lambda x, y, t, u, p: 0.0 * u, FVM.Dirichlet)
    rng = np.random.default_rng(seed)
    N = tri.num_points()
    true = np.log(0.1) + 0.3 * rng.standard_normal(N)
    u0 = np.sin(np.pi * tri.x) * np.sin(np.pi * tri.y) + 0.2 * rng.random(N)
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
