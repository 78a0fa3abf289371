"""The abundance-fit gradients (dispatch_amd.fit) against central finite
differences of the plain FP64 formal solution: the four log-abundance offsets with
h = 1e-3 dex to 1e-3 relative, and directional derivatives in T, on a reduced box
of the fit workload's generator."""
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
    """Directional derivatives of the loss in T against central differences of the
    FP64 loss (the ghost temperature 2 T(1) - T(2) moving with T):

    * a uniform 1 K shift of every cell, where the cells' contributions add: to
      1e-3 relative;
    * a random per-cell direction v (unit normal, 1 K rms), where they nearly cancel
      (sum |dL/dT_i v_i| is about 25 times |<dL/dT, v>| on this box) and the FP64
      difference itself moves by 3e-3 of its value between steps of 0.25 and 2 K
      (the perturbed opacities are rounded to FP32 as the forward's are): to 1e-3
      of the scale of the terms, sum |dL/dT_i v_i|.
    (Single cells are a poorer check still: the loss is piecewise smooth through
    the upwind clamp.)"""
    p = problem
    d0 = np.zeros(4)
    print("fit temperature gradient", flush=True)
    loss, _, gT = p.loss_and_grad(d0, want_T=True)
    g = torch.Generator(device="cuda").manual_seed(11)
    T0 = p.box.T.clone()
    Tg0 = p.box.T_ghost_bot.clone()

    def fd(v, eps):
        vals = []
        for sgn in (1, -1):
            p.box.T.copy_(T0 + sgn * eps * v)
            p.box.T_ghost_bot.copy_(2 * p.box.T[0] - p.box.T[1])
            vals.append(float(p.loss_of_flux(p.flux_fp64(d0))))
        p.box.T.copy_(T0)
        p.box.T_ghost_bot.copy_(Tg0)
        return (vals[0] - vals[1]) / (2 * eps)

    u = torch.ones_like(T0)
    want = float(gT.sum())
    for eps in (1.0, 0.25):
        got = fd(u, eps)
        assert abs(want - got) <= 1e-3 * abs(got), ("uniform", eps, want, got)
    v = torch.randn(T0.shape, generator=g, dtype=torch.float64, device="cuda")
    want = float((gT * v).sum())
    scale = float((gT * v).abs().sum())
    for eps in (1.0, 0.25):
        got = fd(v, eps)
        assert abs(want - got) <= 1e-3 * scale, ("random", eps, want, got, scale)
