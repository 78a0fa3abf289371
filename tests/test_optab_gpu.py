"""GPU parity of the sampled background table (m3drt_optab_plan_create,
m3drt_optab_interpolate; HOP/optab_mod.f90:127-287) against the reference's
compiled optab_mod (tests/golden/hop_optab_sample.npz) and the oracle:

- the node slopes are the reference's bit for bit;
- totals and scattering parts at the reference's query cells are its values bit
  for bit, except where the device and host FP64 logarithms of T or ne round
  differently and move the FP32 Hermite coefficients by an ulp (then within 1e-6);
- on a larger random set of cells and frequencies the kernel equals the oracle
  under the same rule.
"""
import os

import numpy as np
import pytest
import torch

from dispatch_amd import m3drt
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hop_optab_sample.npz")


def dev(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a, dt)).cuda()


def plan_of(ctx, g):
    F = lambda a: dev(np.ascontiguousarray(np.transpose(a, (2, 1, 0))), np.float32)   # (nne, ntemp, nnu)
    return m3drt.OptabPlan(ctx, dev(g["in_lam"], np.float64), F(g["in_val"]), F(g["in_scat"]), g["in_mu"],
                           g["in_rot"], g["in_c0"], g["in_dc"])


def compare(got, want):
    same = got == want
    assert same.mean() > 0.999
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
    assert rel.max() < 1e-6


def test_slopes_and_reference_queries(ctx):
    g = np.load(GOLD)
    p = plan_of(ctx, g)
    sl = [s.cpu().numpy().transpose(2, 1, 0) for s in p.slopes()]
    for name, s in zip(("dl", "dt", "de", "sl", "st", "se"), sl):
        assert np.array_equal(s, g[name]), name
    nc, nf = g["in_temp"].size, g["in_lam_q"].size
    back = torch.empty((nf, nc), dtype=torch.float32, device="cuda")
    scat = torch.empty_like(back)
    p.interpolate(dev(g["in_temp"], np.float64), dev(g["in_ne"], np.float64), dev(g["in_chi500"], np.float32),
                  dev(g["in_lam_q"], np.float64), back, scat)
    torch.cuda.synchronize()
    compare(back.cpu().numpy().T, g["back"])
    compare(scat.cpu().numpy().T, g["scat"])
    p.close()


def test_random_queries_match_oracle(ctx):
    g = np.load(GOLD)
    rng = np.random.default_rng(4)
    mu, rot, c0, dc = g["in_mu"], g["in_rot"], g["in_c0"], g["in_dc"]
    nnu, nt, ne_ = g["in_val"].shape
    nc, nf = 20000, 300
    p1 = c0[0] + rng.uniform(-0.5, nt - 0.5, nc) * dc[0]
    p2 = c0[1] + rng.uniform(-0.5, ne_ - 0.5, nc) * dc[1]
    d = rot @ np.vstack([p1, p2])
    temp, ne = np.exp(d[0] + mu[0]), np.exp(d[1] + mu[1])
    chi500 = (10 ** rng.uniform(-9, -5, nc)).astype(np.float32)
    lq = np.sort(rng.uniform(g["in_lam"][0] - 50, g["in_lam"][-1] + 50, nf))
    wb, ws, _ = O.optab_interpolate(g["in_lam"], g["in_val"], g["in_scat"], mu, rot, c0, dc, temp, ne, chi500, lq)
    p = plan_of(ctx, g)
    back = torch.empty((nf, nc), dtype=torch.float32, device="cuda")
    scat = torch.empty_like(back)
    p.interpolate(dev(temp, np.float64), dev(ne, np.float64), dev(chi500, np.float32), dev(lq, np.float64), back, scat)
    torch.cuda.synchronize()
    compare(back.cpu().numpy().T, wb)
    compare(scat.cpu().numpy().T, ws)
    p.close()
