"""The abundance-fit gradients (dispatch_amd.fit) against central finite
differences of the plain FP64 formal solution: the four log-abundance offsets with
h = 1e-3 dex and the temperature of the most sensitive cells, to 1e-3 relative, on
a reduced box of the fit workload's generator."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem(ctx):
    from dispatch_amd.fit import FitProblem
    p = FitProblem(ctx, nx=16, ny=12, nz=40, nf=96, seed=5)
    p.make_target()
    return p


def test_abundance_gradient_matches_fp64_differences(problem):
    p = problem
    d0 = np.array([0.05, -0.03, 0.0, 0.02])
    print("fit abundance gradient, box 16x12x40, 96 f", flush=True)
    loss, g, _ = p.loss_and_grad(d0)
    h = 1e-3
    for e in range(4):
        dp, dm = d0.copy(), d0.copy()
        dp[e] += h
        dm[e] -= h
        fd = (float(p.loss_of_flux(p.flux_fp64(dp))) - float(p.loss_of_flux(p.flux_fp64(dm)))) / (2 * h)
        assert abs(float(g[e]) - fd) <= 1e-3 * abs(fd) + 1e-6 * loss, (e, float(g[e]), fd)


def test_temperature_gradient_matches_fp64_differences(problem):
    """Along a random per-cell direction v (unit normal, scaled to 1 K rms): the
    central difference of the FP64 loss in T + eps v against <dL/dT, v>.  (Single
    cells are a poorer check: the loss is piecewise smooth through the upwind clamp,
    and a one-cell FP64 difference moves by up to 0.5 % with the step.)"""
    p = problem
    d0 = np.zeros(4)
    print("fit temperature gradient", flush=True)
    loss, _, gT = p.loss_and_grad(d0, want_T=True)
    g = torch.Generator(device="cuda").manual_seed(11)
    v = torch.randn(p.box.T.shape, generator=g, dtype=torch.float64, device="cuda")
    T0 = p.box.T.clone()
    Tg0 = p.box.T_ghost_bot.clone()
    want = float((gT * v).sum())
    for eps in (1.0, 0.25):
        vals = []
        for sgn in (1, -1):
            p.box.T.copy_(T0 + sgn * eps * v)
            p.box.T_ghost_bot.copy_(2 * p.box.T[0] - p.box.T[1])
            vals.append(float(p.loss_of_flux(p.flux_fp64(d0))))
        fd = (vals[0] - vals[1]) / (2 * eps)
        p.box.T.copy_(T0)
        p.box.T_ghost_bot.copy_(Tg0)
        assert abs(want - fd) <= 1e-3 * abs(fd), (eps, want, fd)
