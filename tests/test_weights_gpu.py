"""The FP32 step weights every sweep kernel inlines (dev::lin_weights,
dev::parab_moments), run on the GPU through m3drt_linear_weights and
m3drt_parab_moments, against the FP64 closed forms:

  E = e^-t, u0 = -expm1(-t), Psi = 1 - u0/t, W = u0 - Psi, Esc = 1 - Psi,
  U0 = u0, U1 = t - u0, q = t - 2 U1/t (Taylor series of Psi and q near t = 0).

The reference states its weights within 4 FP32 ulp over -2 <= t <= 90
(SRC/rt_solvers_mod.f90:772-774, 857-859; E to the clamp at 87).  Its own FP32
arithmetic (the oracle's restatement, bit-exact with the compiled exp_neg) reaches
4.8 ulp in W near t = 0.87, so the bound per output is max(4, the reference
arithmetic's worst on the same grid) plus one ulp for the hardware reciprocal and
exponential.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

F = np.float32


def grid():
    t = np.concatenate([np.linspace(-2, 90, 600001), np.linspace(-1e-3, 1e-3, 4001),
                        np.linspace(0.99, 1.01, 4001), np.linspace(-1.01, -0.99, 4001),
                        np.linspace(1.99, 2.01, 4001), np.linspace(-2.0, -1.99, 2001),
                        np.linspace(86.9, 87.1, 2001)])
    t = t.astype(F)
    return t[t != 0]


def exact(td):
    E = np.exp(-td)
    u0 = -np.expm1(-td)
    small = np.abs(td) < 1e-2
    with np.errstate(all="ignore"):
        psi = np.where(small, td / 2 - td ** 2 / 6 + td ** 3 / 24 - td ** 4 / 120 + td ** 5 / 720, 1 - u0 / td)
        U1 = td - u0
        q = np.where(small, td ** 2 / 3 - td ** 3 / 12 + td ** 4 / 60 - td ** 5 / 360, td - 2 * U1 / td)
    return dict(E=E, u0=u0, psi=psi, W=u0 - psi, esc=1 - psi, U1=U1, q=q)


def ulps(got, ex):
    sp = np.spacing(np.abs(ex).astype(F)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ex) / sp


def bound(ref_worst):
    return max(4.0, ref_worst) + 1.0


def test_linear_weights_within_the_reference_bounds(ctx):
    t = grid()
    td = t.astype(np.float64)
    x = exact(td)
    g = [a.cpu().numpy() for a in ctx.linear_weights(torch.from_numpy(t).cuda())]
    r = O.linear_weights(t)
    m87 = td <= 87
    report = {}
    for k, (name, ex) in enumerate([("E", x["E"]), ("W", x["W"]), ("Psi", x["psi"]), ("Esc", x["esc"])]):
        sel = m87 if name == "E" else np.ones_like(m87)
        ug, ur = ulps(g[k][sel], ex[sel]).max(), ulps(r[k][sel], ex[sel]).max()
        report[name] = (ug, ur)
        assert ug <= bound(ur), (name, ug, ur)
    # past the clamp the exponential holds e^-87 (below every quantity it multiplies)
    assert np.all(g[0][td > 87] == g[0][np.argmin(np.abs(td - 87))])
    print("linear_weights max ulp (gpu, reference arithmetic):", report)


def test_parab_moments_within_the_reference_bounds(ctx):
    t = grid()
    t = t[t <= 87]
    td = t.astype(np.float64)
    x = exact(td)
    g = [a.cpu().numpy() for a in ctx.parab_moments(torch.from_numpy(t).cuda())]
    r = O.parab_moments(t)
    report = {}
    for k, (name, ex) in enumerate([("E", x["E"]), ("U0", x["u0"]), ("U1", x["U1"]), ("q", x["q"])]):
        ug, ur = ulps(g[k], ex).max(), ulps(r[k], ex).max()
        report[name] = (ug, ur)
        assert ug <= bound(ur), (name, ug, ur)
    print("parab_moments max ulp (gpu, reference arithmetic):", report)


def test_weights_of_empty_input(ctx):
    t = torch.empty(0, dtype=torch.float32, device="cuda")
    assert all(a.numel() == 0 for a in ctx.linear_weights(t))
