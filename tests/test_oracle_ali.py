"""Pins the oracle's restatement of add_bound_bound_opacities with add_largest
(SRC/rt_patch_mod.f90:2689-2831, 3150-3169), of the preconditioned 'largest' rate
integrands (:3062-3098, with other_contribution :3174-3187) and of
bound_free_rates (:3193-3259) by what the mathematics fixes:

* line opacity: eta/chi = b_j/b_i per point (1 in LTE); the split tracking holds the
  largest-magnitude term and its index, and largest + rest is the total, against a
  line-by-line recount; wphi agrees with the rates' own normalisation;
* split rates: with solver outputs consistent with the tracked opacities (Irem = I -
  Psi eta, escape = 1 - Psi chi) and beta = beta_ref, preconditioning leaves the net
  rate n_i R_ij - n_j R_ji unchanged (the Rybicki & Hummer construction); where Psi
  is zero, or every psi flag is in force, the integrands are the unpreconditioned
  ones, bit for bit;
* bf rates: LTE detailed balance (J = B gives R_ij = R_ji), radiative_cut, and the
  preconditioning correction linear in PsiJ and equal for both directions when
  b_i = b_j.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rates_cases import (level_products, make_equal_pair_case, make_rates_case, rate_coefficients,
                         split_solver_outputs)
from test_oracle_rates import magnitude

WMU = 1.0 / 24


def zeros_split(nf, ncell):
    z = np.zeros((nf, ncell), np.float32)
    zi = np.zeros((nf, ncell), np.int32)
    return dict(split_opac1=z, split_opac2=z, split_chi1=z, split_chi2=z, split_id=zi, split_chi_id=zi)


def test_line_opacity_ratio_and_lte():
    d = make_rates_case(seed=21, ncell=40, nf=150, nline=1)
    nf, ncell = d["Bnu"].shape
    z = np.zeros((nf, ncell), np.float32)
    op = O.bb_opacity(d, z, z)
    li, lj = d["lev_i"][0] - 1, d["lev_j"][0] - 1
    m = op["xhat"] > 0
    ratio = (d["b"][lj].astype(np.float32) / d["b"][li].astype(np.float32))[None, :] * np.ones((nf, 1))
    np.testing.assert_allclose(op["ehat"][m] / op["xhat"][m], ratio[m], rtol=3e-7)
    lte = O.bb_opacity(d, z, z, lte=True)
    np.testing.assert_array_equal(lte["ehat"], lte["xhat"])
    # accumulates
    twice = O.bb_opacity(d, op["xhat"], op["ehat"])
    np.testing.assert_allclose(twice["xhat"], 2 * op["xhat"], rtol=1e-6)


@pytest.mark.parametrize("seed", [22, 23])
def test_split_tracking_recount(seed):
    d = make_rates_case(seed=seed, ncell=30, nf=160, nline=6)
    nline, nf = d["q"].shape
    ncell = d["b"].shape[1]
    z = np.zeros((nf, ncell), np.float32)
    full = O.bb_opacity(d, z, z, split=zeros_split(nf, ncell), split_offset=10)
    terms_e, terms_c = [], []
    for l in range(nline):
        dl = {k: (v[l:l + 1] if k in ("iblue", "ired", "lev_i", "lev_j", "Bij", "lower_off", "q", "wq", "adamp")
                  else v) for k, v in d.items()}
        one = O.bb_opacity(dl, z, z)
        terms_e.append(one["ehat"])
        terms_c.append(one["xhat"])
    for terms, big, rest, idx in ((terms_e, "split_opac1", "split_opac2", "split_id"),
                                  (terms_c, "split_chi1", "split_chi2", "split_chi_id")):
        t = np.stack(terms)                                   # (nline, nf, ncell)
        first_max = np.argmax(np.abs(t), axis=0)              # first of equal maxima, as the strict '>'
        any_term = np.abs(t).max(axis=0) > 0
        np.testing.assert_array_equal(full[big][any_term],
                                      np.take_along_axis(t, first_max[None], 0)[0][any_term])
        np.testing.assert_array_equal(full[idx][any_term], (first_max + 11)[any_term])
        np.testing.assert_array_equal(full[idx][~any_term], 0)
        tot = t.astype(np.float64).sum(0)
        np.testing.assert_allclose(full[big] + full[rest].astype(np.float64), tot,
                                   rtol=1e-6, atol=1e-6 * np.abs(tot).max())
        # and the total is what xhat / ehat hold
        np.testing.assert_allclose(full["ehat" if big == "split_opac1" else "xhat"], tot,
                                   rtol=1e-6, atol=1e-6 * np.abs(tot).max())


def test_wphi_matches_rates_normalisation():
    d = make_rates_case(seed=24, ncell=50, nf=200, nline=4)
    nline, nf = d["q"].shape
    ncell = d["b"].shape[1]
    z = np.zeros((nf, ncell), np.float32)
    w0 = np.zeros((nline, ncell), np.float32)
    op = O.bb_opacity(d, z, z, wmu=1.0, wphi_line=np.ones(nline, np.int32), wphi=w0)
    zg = np.zeros((nline, ncell), np.float32)
    plain = O.bb_rates(d, WMU, 1.0, zg, zg)
    norm = O.bb_rates(d, WMU, 1.0, zg, zg, own_norm=True)
    sw = plain[0] / norm[0]  # = sum phi wq (real(4) weights) * wtot
    np.testing.assert_allclose(op["wphi"], sw, rtol=2e-6)
    off = O.bb_opacity(d, z, z, wmu=1.0, wphi_line=np.zeros(nline, np.int32), wphi=w0)
    assert np.all(off["wphi"] == 0)


def _split_case(seed):
    d = make_equal_pair_case(seed=seed)
    nline, nf = d["q"].shape
    ncell = d["b"].shape[1]
    z = np.zeros((nf, ncell), np.float32)
    opac = O.bb_opacity(d, z, z, split=zeros_split(nf, ncell), split_offset=3)
    chi = opac["xhat"] - opac["ehat"] * d["beta"]
    rng = np.random.default_rng(seed + 100)
    Psi = np.where(chi > 0, rng.uniform(0.05, 0.8, chi.shape) / np.where(chi > 0, chi, 1.0), 0.0).astype(np.float32)
    d["Psi"] = Psi
    return d, split_solver_outputs(d, opac, Psi)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_split_rates_keep_the_net_rate(seed):
    d, sp = _split_case(seed)
    nline = d["q"].shape[0]
    ncell = d["b"].shape[1]
    zg = np.zeros((nline, ncell), np.float32)
    g1 = O.bb_rates(d, WMU, 1.0, zg, zg, precond=True, largest=True, split=sp, split_offset=3)
    g0 = O.bb_rates(d, WMU, 1.0, zg, zg, largest=True)
    ni, nj = level_products(d)
    R0, R1 = rate_coefficients(d, *g0), rate_coefficients(d, *g1)
    net0 = ni * R0[0] - nj * R0[1]
    net1 = ni * R1[0] - nj * R1[1]
    scale = magnitude(d, *g0) + magnitude(d, *g1)
    assert np.max(np.abs(net1 - net0) / scale) < 2e-6
    assert np.median(np.abs(R1[0] - R0[0]) / np.abs(R0[0])) > 1e-2


def test_split_rates_reduce_without_psi():
    d, sp = _split_case(4)
    nline = d["q"].shape[0]
    ncell = d["b"].shape[1]
    zg = np.zeros((nline, ncell), np.float32)
    plain = O.bb_rates(d, WMU, 1.0, zg, zg, largest=True)
    d0 = dict(d, Psi=np.zeros_like(d["Psi"]))
    for a, b in zip(O.bb_rates(d0, WMU, 1.0, zg, zg, precond=True, largest=True, split=sp, split_offset=3), plain):
        np.testing.assert_array_equal(a, b)
    # every psi flag in force: the unpreconditioned integrands, and nothing raised
    ntrans = np.arange(1, nline + 1, dtype=np.int32) * 7
    npw = (int(ntrans.max()) - 1) // 32 + 1
    allon = dict(sp, ntrans=ntrans, psi_flags_in=np.full((npw, ncell), -1, np.int32))
    gij, gji, out = O.bb_rates(d, WMU, 1.0, zg, zg, precond=True, largest=True, split=allon, split_offset=3)
    np.testing.assert_array_equal(gij, plain[0])
    np.testing.assert_array_equal(gji, plain[1])
    assert np.all(out == 0)
    # no flag in force, Psi zero everywhere: every (line, cell) raises its bit
    none = dict(sp, ntrans=ntrans, psi_flags_in=np.zeros((npw, ncell), np.int32))
    _, _, out = O.bb_rates(d0, WMU, 1.0, zg, zg, precond=True, largest=True, split=none, split_offset=3)
    want = np.zeros((npw, ncell), np.int64)
    for t in ntrans:
        want[(t - 1) // 32] |= 1 << ((t - 1) % 32)
    np.testing.assert_array_equal(out.astype(np.int64) & 0xFFFFFFFF, want & 0xFFFFFFFF)


def bf_case(seed, ncell=60, nf=200, ncont=3):
    rng = np.random.default_rng(seed)
    nline = 4
    nk = 4
    ev = np.array([0.0, 1.0, 2.0, 7.0])
    T = rng.uniform(3000.0, 9000.0, ncell)
    e_kT = synth_ekt(T)
    nstar = 10 ** rng.uniform(6, 10, (nk, ncell))
    b = rng.uniform(0.5, 2.0, (nk, ncell))
    lev_i = np.array([1, 2, 3], np.int32)[:ncont]
    lev_j = np.full(ncont, 4, np.int32)
    iblue = np.array([1, 40, 90], np.int32)[:ncont]
    ired = np.array([120, 200, 150], np.int32)[:ncont]
    nu = np.linspace(1.5e15, 2.5e15, nf)
    wnu = np.zeros((ncont, nf))
    alpha = np.zeros((ncont, nf), np.float32)
    for j in range(ncont):
        s = slice(iblue[j] - 1, ired[j])
        wnu[j, s] = np.gradient(nu)[s]
        alpha[j, s] = (1e-18 * (nu[0] / nu[s]) ** 3).astype(np.float32)
    beta = np.exp(-6.62607015e-27 * nu[:, None] / (1.380649e-16 * T[None, :])).astype(np.float32)
    Bnu = (beta.astype(np.float64) / (1.0 - beta)).astype(np.float32)
    return dict(iblue=iblue, ired=ired, lev_i=lev_i, lev_j=lev_j, active=np.ones(ncont, np.int32), ev=ev, wnu=wnu,
                alpha=alpha, twohnu3_c2=(1.4745e-47 * nu ** 3).astype(np.float32), e_kT=e_kT, nstar=nstar, b=b,
                Jnu=Bnu, beta=beta, PsiJ=np.zeros((nf, ncell), np.float32)), nline


def synth_ekt(T):
    return 11604.518 / np.asarray(T, np.float64)


def test_bf_rates_detailed_balance_and_cut():
    d, nline = bf_case(31)
    ncont, nf = d["alpha"].shape
    ncell = d["nstar"].shape[1]
    z = np.zeros((nline + ncont, ncell), np.float32)
    gij, gji = O.bf_rates(d, nline, z, z)
    assert np.all(gij[:nline] == 0) and np.all(gji[:nline] == 0)
    live = gij[nline:] > 0
    np.testing.assert_allclose(gji[nline:][live], gij[nline:][live], rtol=2e-5)
    # radiative_cut: continua whose threshold over kT exceeds cut_x are skipped
    cut = O.bf_rates(d, nline, z, z, cut_x=15.0)
    dev = d["ev"][d["lev_j"] - 1] - d["ev"][d["lev_i"] - 1]
    skip = dev[:, None] * d["e_kT"][None, :] > 15.0
    assert skip.any() and (~skip).any()
    assert np.all(cut[0][nline:][skip] == 0)
    np.testing.assert_array_equal(cut[0][nline:][~skip], gij[nline:][~skip])
    # an inactive continuum adds nothing
    off = O.bf_rates(dict(d, active=np.array([1, 0, 1], np.int32)), nline, z, z)
    assert np.all(off[0][nline + 1] == 0)


def test_bf_rates_preconditioning():
    d, nline = bf_case(32)
    ncont, nf = d["alpha"].shape
    ncell = d["nstar"].shape[1]
    z = np.zeros((nline + ncont, ncell), np.float32)
    rng = np.random.default_rng(5)
    P = rng.uniform(0.0, 1e-3, (nf, ncell)).astype(np.float32)
    base = O.bf_rates(d, nline, z, z, precond=True)
    one = O.bf_rates(dict(d, PsiJ=P), nline, z, z, precond=True)
    two = O.bf_rates(dict(d, PsiJ=2 * P), nline, z, z, precond=True)
    for k in range(2):
        d1 = one[k][nline:].astype(np.float64) - base[k][nline:]
        d2 = two[k][nline:].astype(np.float64) - base[k][nline:]
        np.testing.assert_allclose(d2, 2 * d1, rtol=1e-3, atol=1e-6 * np.abs(base[k]).max())
        assert np.all(d1 <= 0)
    # with b_i = b_j both directions lose the same amount
    bb = d["b"].copy()
    bb[:] = 1.3
    b1 = O.bf_rates(dict(d, b=bb), nline, z, z, precond=True)
    b2 = O.bf_rates(dict(d, b=bb, PsiJ=P), nline, z, z, precond=True)
    np.testing.assert_allclose(b2[0].astype(np.float64) - b1[0], b2[1].astype(np.float64) - b1[1],
                               rtol=1e-4, atol=1e-6 * np.abs(b1[0]).max())
