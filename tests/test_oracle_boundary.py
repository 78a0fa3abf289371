"""Pins the oracle's boundary planes and LTE background factors
(SRC/rt_patch_mod.f90:2004-2023, 3270-3374):

- the PCHIP bottom ghost equals the reference's compiled pchip_mono at one plane
  below the first three (tests/golden/rt_pchip_ghost.npz, from
  oracle/ref_harness/rt_golden.f90) bit for bit, and its closed form 4 y1 - 3 y2
  (the first interval's Hermite cubic with the slope 2 dy1/dx1 at node 1 and the
  never-formed slope 0 at node 2, evaluated at t = -1) to FP32 rounding;
- beta = exp(-h nu/kT) clamped at the smallest normal FP32 number and
  Bnu = beta/(1 - beta) = 1/(exp(h nu/kT) - 1), against FP64 with the reference's
  constants h = 6.626176e-27, k = 1.380662e-16 (HOP/const_mod.f90:39-43);
- the LTE ghost plane: chi = chi(first plane) g500/chi500(first plane), S = I =
  Bnu(T_ghost).
"""
import os

import numpy as np

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
H_K = 6.626176e-27 / 1.380662e-16


def test_pchip_ghost_is_the_reference_build():
    g = np.load(os.path.join(GOLD, "rt_pchip_ghost.npz"))
    y = g["in_y"]
    got = O.pchip_ghost(y[:, 0], y[:, 1], y[:, 2], float(g["in_ds3"]))
    assert np.array_equal(got, g["ynew"])
    cf = 4 * y[:, 0].astype(np.float64) - 3 * y[:, 1].astype(np.float64)
    scale = np.abs(y[:, :2].astype(np.float64)).max(axis=1) + 1e-300
    assert np.max(np.abs(got - cf) / scale) < 3e-6


def test_pchip_ghost_other_cell_sizes():
    rng = np.random.default_rng(3)
    y = (10 ** rng.uniform(-9, -3, (3, 500))).astype(np.float32)
    for ds3 in (1.0, 3.7e5, 1e7):
        got = O.pchip_ghost(y[0], y[1], y[2], ds3)
        cf = 4 * y[0].astype(np.float64) - 3 * y[1].astype(np.float64)
        assert np.max(np.abs(got - cf) / (np.abs(y[:2]).max(axis=0))) < 3e-6


def test_lte_background_closed_form():
    rng = np.random.default_rng(5)
    nu = np.concatenate([np.geomspace(1e12, 3e16, 300), [5e16, 1e17]])
    T = np.concatenate([rng.uniform(2500, 30000, 200), [3000.0, 1e5]])
    beta, B = O.lte_background(nu, T)
    x = H_K * nu[:, None] / T.astype(np.float32).astype(np.float64)[None, :]
    b64 = np.exp(-x)
    tiny = np.finfo(np.float32).tiny
    clamp = b64 < tiny
    assert np.all(beta[clamp] == tiny)
    ok = ~clamp
    # h_kT and hv_kT are rounded to FP32 before the exponential: |x| ulp-scale error
    assert np.max(np.abs(beta[ok] / b64[ok] - 1) / np.maximum(x[ok], 1)) < 3e-7
    with np.errstate(over="ignore"):
        B64 = 1 / np.expm1(x)
    # B = beta/(1 - beta) in FP32 carries beta's rounding amplified by 1/(1 - beta)
    bound = 3e-7 * np.maximum(x[ok], 1) + 1.2e-7 / (1 - b64[ok])
    assert np.all(np.abs(B[ok] / B64[ok] - 1) < bound)
    assert clamp.any() and ok.any()


def test_lte_ghost_plane():
    rng = np.random.default_rng(7)
    nf, nxy = 40, 300
    nu = np.linspace(5.9e14, 6.0e14, nf)
    chi1 = (10 ** rng.uniform(-8, -5, (nf, nxy))).astype(np.float32)
    g500 = (10 ** rng.uniform(-7, -6, nxy)).astype(np.float32)
    c500 = (10 ** rng.uniform(-7.3, -6.3, nxy)).astype(np.float32)
    T = rng.uniform(6000, 9000, nxy).astype(np.float32)
    gchi, gS, gI = O.lte_ghost(nu, chi1, g500, c500, T)
    r500 = (g500 / c500).astype(np.float32)
    assert np.array_equal(gchi, (chi1 * r500[None, :]).astype(np.float32))
    beta, B = O.lte_background(nu, T.astype(np.float64))
    assert np.array_equal(gS, B) and np.array_equal(gI, gS)
