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
    p = problem
    d0 = np.zeros(4)
    print("fit temperature gradient", flush=True)
    loss, _, gT = p.loss_and_grad(d0, want_T=True)
    flat = gT.abs().flatten()
    cells = torch.topk(flat, 3).indices.tolist()
    T0 = p.box.T.clone()
    Tg0 = p.box.T_ghost_bot.clone()
    h = 2.0
    for c in cells:
        z, rem = divmod(c, p.box.ny * p.box.nx)
        y, x = divmod(rem, p.box.nx)
        vals = []
        for sgn in (1, -1):
            p.box.T.copy_(T0)
            p.box.T[z, y, x] += sgn * h
            p.box.T_ghost_bot.copy_(2 * p.box.T[0] - p.box.T[1])
            vals.append(float(p.loss_of_flux(p.flux_fp64(d0))))
        fd = (vals[0] - vals[1]) / (2 * h)
        p.box.T.copy_(T0)
        p.box.T_ghost_bot.copy_(Tg0)
        assert abs(float(gT[z, y, x]) - fd) <= 1e-3 * abs(fd), ((z, y, x), float(gT[z, y, x]), fd)
