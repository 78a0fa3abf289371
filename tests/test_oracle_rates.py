"""Pins the oracle's restatement of bound_bound_rates (SRC/rt_patch_mod.f90:2852-3142)
by what its mathematics fixes, not by retyping it:

* closed form: Bnu and Inu flat across a line, wnu = nu0 wq and the line's own
  normalisation give th = A_ij (wmu/wtot) nu0 Bnu / n_ref, whatever the profile;
* detailed balance: for Inu = Bnu without preconditioning the deviation integral
  is zero ('largest' exactly, 'original' to rounding), and the rates then satisfy
  n*_i R_ij = n*_j R_ji;
* the preconditioning corrections cancel in the net rate n_i R_ij - n_j R_ji at the
  populations they were formed with (reference comment at :3112-3118);
* the own normalisation is invariant under a common scaling of wnu and wq, and
  scales as 1/wtot;
* Kahan and plain sums agree to FP32 rounding of the sums.
The profile itself (QLG-bot) is pinned bit for bit against the reference's compiled
qlg_mod in test_golden_hop.py.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rates_cases import level_products, make_rates_case, rate_coefficients

WMU = 1.0 / 24


def run(d, **kw):
    z = np.zeros(d["adamp"].shape, np.float32)
    return O.bb_rates(d, WMU, kw.pop("wtot", 1.0), z, z, **kw)


def n_ref(d):
    li, lj = d["lev_i"] - 1, d["lev_j"] - 1
    g = d["g"]
    beta_ref = d["nstar"][lj] * g[li][:, None] / (d["nstar"][li] * g[lj][:, None])
    return beta_ref / (1.0 - beta_ref)


def magnitude(d, gij, gji):
    """Sum of the magnitudes of the terms of n_i R_ij and n_j R_ji."""
    li, lj = d["lev_i"] - 1, d["lev_j"] - 1
    g = d["g"]
    nr = n_ref(d)
    th, dev = np.abs(gij.astype(np.float64)), np.abs(gji.astype(np.float64))
    ni, nj = level_products(d)
    return ni * (nr * th + dev) + nj * (g[li] / g[lj])[:, None] * ((nr + 1) * th + dev)


@pytest.mark.parametrize("kahan", [False, True])
def test_flat_line_closed_form(kahan):
    d = make_rates_case(seed=3, flat=True)
    gij, gji = run(d, own_norm=True, largest=True, kahan=kahan, wtot=0.75)
    li, lj = d["lev_i"] - 1, d["lev_j"] - 1
    nu0 = d["wnu"].sum(1) / d["wq"].sum(1)
    th = d["Aij"][:, None] * (WMU / 0.75) * nu0[:, None] * d["Bnu"][0][None, :].astype(np.float64) / n_ref(d)
    np.testing.assert_allclose(gij, th, rtol=2e-5)
    assert np.all(gji == 0.0)


@pytest.mark.parametrize("seed", [0, 1])
def test_detailed_balance(seed):
    d = make_rates_case(seed=seed)
    d["Inu"] = d["Bnu"]
    g_l = run(d, largest=True)
    assert np.all(g_l[1] == 0.0)
    g_o = run(d)
    nr = n_ref(d)
    # dev = sum_I - sum_A/inv_nref: rounding of the FP32 sums only
    np.testing.assert_allclose(g_o[1] / (nr * g_o[0]), 0.0, atol=3e-6)
    Rij, Rji = rate_coefficients(d, *g_l)
    li, lj = d["lev_i"] - 1, d["lev_j"] - 1
    np.testing.assert_allclose(d["nstar"][li] * Rij, d["nstar"][lj] * Rji, rtol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2, 4])
def test_preconditioning_cancels_in_net_rate(seed):
    d = make_rates_case(seed=seed)
    g0, g1 = run(d), run(d, precond=True)
    ni, nj = level_products(d)
    R0, R1 = rate_coefficients(d, *g0), rate_coefficients(d, *g1)
    net0 = ni * R0[0] - nj * R0[1]
    net1 = ni * R1[0] - nj * R1[1]
    scale = magnitude(d, *g0) + magnitude(d, *g1)
    assert np.max(np.abs(net1 - net0) / scale) < 1e-6
    # and the correction is not vacuous: it moves R_ij by more than 1 % in most cells
    assert np.median(np.abs(R1[0] - R0[0]) / np.abs(R0[0])) > 1e-2


def test_own_normalisation_scalings():
    d = make_rates_case(seed=5)
    base = run(d, own_norm=True)
    d2 = dict(d, wnu=d["wnu"] * 8.0, wq=d["wq"] * 8.0)
    sc = run(d2, own_norm=True)
    for a, b in zip(base, sc):
        np.testing.assert_allclose(a, b, rtol=2e-6, atol=1e-6 * np.abs(a).max())
    half = run(d, own_norm=True, wtot=0.5)
    for a, b in zip(base, half):
        np.testing.assert_allclose(b, 2.0 * a, rtol=2e-6, atol=1e-6 * np.abs(a).max())
    plain = run(d)
    # normalised = plain / (sum phi wq * wtot): the ratio is one number per (line, cell)
    r = plain[0] / base[0]
    np.testing.assert_allclose(plain[1] / r, base[1], rtol=1e-5, atol=1e-5 * np.abs(base[1]).max())


def test_kahan_matches_plain_sums():
    d = make_rates_case(seed=6, nf=400)
    for kw in (dict(), dict(precond=True), dict(largest=True)):
        a, b = run(d, **kw), run(d, kahan=True, **kw)
        for x, y in zip(a, b):
            np.testing.assert_allclose(x, y, rtol=1e-4, atol=1e-5 * np.abs(y).max())


def test_accumulates_into_existing_rates():
    d = make_rates_case(seed=7)
    one = run(d)
    z0 = np.full(d["adamp"].shape, 3.0, np.float32)
    acc = O.bb_rates(d, WMU, 1.0, z0, -z0)
    np.testing.assert_allclose(acc[0], one[0] + 3.0, rtol=1e-6)
    np.testing.assert_allclose(acc[1], one[1] - 3.0, rtol=1e-6)
