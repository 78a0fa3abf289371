"""The oracle of the batched 1D synthesis (oracle/m3d_b1d.c) pinned by exact
answers, and the workload generator's and the binding's host logic.

* A line-free atmosphere whose source function is linear in tau500 has the
  Eddington-Barbier emergent intensity a + b mu, which the linear step reproduces
  exactly (the deepest point at tau500 = 100 leaves e^-100 of the bottom value).
* An isothermal atmosphere, lines or not, emits I = B at every mu.
* The line opacity at a point equals s K(a, v) with K the oracle's qlg_bot_sweep
  (itself bit-exact with the compiled reference), and Re w(v + ia) to QLG-bot's
  stated 5.7e-3 of the peak.
* Outside its window a line of the generated workload stays below 2.5 tol of the
  continuum (the Gaussian core and the Lorentz wing each below tol at the edge).
"""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from dispatch_amd import batch1d as B
from dispatch_amd import m3drt
from oracle import oracle as O

LAM = 500e-7


def one_atm(nd, nf, nmu, nline, dtau, hckT, ivth=None, cont_a=None, cont_b=None, lines=None, dlam=1e-9):
    """Host arrays of a single atmosphere in the batch's field names."""
    mu, w = B.gauss_mu(nmu)
    lines = lines or []
    rec = np.zeros((nd, max(nline, 0), 4), np.float32)
    for l, (off, a, s, reach) in enumerate(lines):
        rec[:, l] = (a, s, reach, 0.0)
    return dict(nd=[nd], nf=[nf], nmu=[nmu], nline=[nline], d_off=[0], l_off=[0], ld_off=[0], mu_off=[0],
                f_off=[0], i_off=[0], lam0=[LAM], dlam=[dlam], dtau=np.asarray(dtau, np.float32),
                hckT=np.asarray(hckT, np.float32),
                ivth=np.asarray(ivth if ivth is not None else np.full(nd, 1e-5), np.float32),
                cont_a=np.asarray(cont_a if cont_a is not None else np.ones(nd), np.float32),
                cont_b=np.asarray(cont_b if cont_b is not None else np.zeros(nd), np.float32),
                line_off=np.asarray([l[0] for l in lines], np.float32),
                line_cl=np.full(len(lines), 2.99792458e10 / LAM, np.float32),
                line_rec=rec.reshape(-1), mu=mu.astype(np.float32), wflux=(2 * w * mu).astype(np.float32))


def test_oracle_eddington_barbier():
    nd, a, b = 600, 0.2, 0.7
    tau = np.concatenate([[0.0], np.logspace(-6, 2, nd - 1)])
    S = a + b * tau
    hckT = LAM * np.log1p(1.0 / S)                      # B(T) = S at lam0
    h = one_atm(nd, 1, 5, 0, np.append(np.diff(tau), 0.0), hckT)
    I, F = O.b1d_atmosphere(h, 0, LAM)
    Sf = 1.0 / np.expm1(h["hckT"].astype(np.float64) / LAM)   # what the float hckT gives
    a_eff = Sf[0]
    mu = h["mu"].astype(np.float64)
    # S is linear in tau to the float rounding of hckT; the linear scheme is exact for it
    np.testing.assert_allclose(I[:, 0], a_eff + b * mu, rtol=2e-6)
    np.testing.assert_allclose(F[0], np.sum(h["wflux"] * I[:, 0]), rtol=1e-12)


def test_oracle_isothermal_is_planck_with_lines():
    nd = 40
    tau = np.logspace(-5, 2, nd)
    hckT = np.full(nd, 2.5e-4)
    lines = [(2e-9, 0.01, 300.0, 5e-9), (6e-9, 0.05, 20.0, 2e-9)]
    h = one_atm(nd, 12, 4, 2, np.append(np.diff(tau), 0.0), hckT, lines=lines)
    I, F = O.b1d_atmosphere(h, 0, LAM)
    lam = LAM + np.arange(12) * 1e-9
    Bv = 1.0 / np.expm1(np.float32(2.5e-4) / lam)
    np.testing.assert_allclose(I, np.broadcast_to(Bv, I.shape), rtol=1e-12)


def test_oracle_line_opacity_is_qlg_bot():
    from scipy.special import wofz
    nd, nf = 2, 400
    a, s = 0.03, 7.0
    off = 200 * 1e-10
    ivth = 1.0 / 2.3e5
    h = one_atm(nd, nf, 3, 1, [1.0, 0.0], [2.5e-4, 2.5e-4], ivth=np.full(nd, ivth), cont_a=np.zeros(nd),
                lines=[(off, a, s, 1.0)], dlam=1e-10)
    _, _, chi = O.b1d_atmosphere(h, 0, LAM, want_chi=True)
    # v of each frequency as the oracle forms it (FP32 offsets), K by qlg_bot_sweep
    xf = (np.arange(nf) * 1e-10).astype(np.float32)
    dl = np.abs(xf - np.float32(off)).astype(np.float64)
    v = dl * (2.99792458e10 / LAM) * ivth
    Kref = np.array([O.qlg_bot(a, np.array([vv], np.float32))[0] for vv in v])
    np.testing.assert_allclose(chi[0], s * Kref, rtol=2e-6, atol=1e-6 * s)
    exact = np.real(wofz(v + 1j * a))
    assert np.max(np.abs(chi[0] / s - exact)) < 5.7e-3 * np.max(exact)


def test_generated_windows_bound_the_tolerance():
    from scipy.special import wofz
    arrays, meta = B.make_batch(B.B1dConfig(natm=4, nd=(10, 20), nf=(50, 80), nline=(20, 40)), device="cpu")
    rec = arrays["line_rec"].view(-1, 4).double().numpy()
    n = len(meta["nd"])
    atm = np.repeat(np.arange(n), meta["nline"].astype(np.int64) * meta["nd"])
    i = np.arange(len(rec)) - meta["ld_off"][atm]
    d = i // meta["nline"][atm]
    l = i - d * meta["nline"][atm]
    dg = meta["d_off"][atm] + d
    lg = meta["l_off"][atm] + l
    ivth = arrays["ivth"].double().numpy()[dg]
    cl = arrays["line_cl"].double().numpy()[lg]
    live = rec[:, 2] > 0
    assert live.any() and (~live).any()
    v_edge = 1.0001 * rec[live, 2] * cl[live] * ivth[live]
    K = np.real(wofz(v_edge + 1j * rec[live, 0]))
    cmin = 0.7 + 0.3 * (500.0 / 500.0) ** 3
    assert np.max(rec[live, 1] * K) <= 2.5 * B.TOL * cmin
    # a skipped line is below the tolerance even at its centre
    assert np.all(rec[~live, 1] <= B.TOL * cmin * (1 + 1e-6))


def test_generator_is_seeded_and_consistent():
    cfg = B.B1dConfig(natm=5, nd=(10, 20), nf=(30, 60), nline=(3, 9))
    a1, m1 = B.make_batch(cfg, device="cpu")
    a2, _ = B.make_batch(cfg, device="cpu")
    for k in a1:
        assert torch.equal(a1[k], a2[k]), k
    for k, dt in m3drt.B1D_DTYPES.items():
        assert a1[k].dtype == dt, k
    m3drt.b1d_check_sizes(a1)
    assert B.updates(m1) == int((m1["nd"].astype(np.int64) * m1["nf"] * m1["nmu"]).sum())
    bad = dict(a1, line_rec=a1["line_rec"][:-4])
    with pytest.raises(ValueError, match="line_rec"):
        m3drt.b1d_check_sizes(bad)
    bad = dict(a1, nmu=torch.full_like(a1["nmu"], 9))
    with pytest.raises(ValueError, match="nmu"):
        m3drt.b1d_check_sizes(bad)


def test_batch_struct_is_built_from_device_tensors_only():
    arrays, _ = B.make_batch(B.B1dConfig(natm=2, nd=(4, 6), nf=(10, 20), nline=(1, 2)), device="cpu")
    with pytest.raises(ValueError, match="device tensor"):
        m3drt._b1d_struct(arrays, LAM)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7])
def test_radau_rule_includes_mu1_and_is_exact_to_degree_2n_minus_2(n):
    x, w = B.radau_mu(n)
    assert x[-1] == 1.0 and np.all(np.diff(x) > 0) and np.all(w > 0)
    for k in range(2 * n - 1):
        assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-13
    if n > 1:
        assert abs(np.sum(w * x ** (2 * n - 1)) - 1.0 / (2 * n)) > 1e-12


def test_config0_shape():
    arrays, meta = B.make_batch(B.CONFIG0, device="cpu")
    assert list(meta["nd"]) == [72] and list(meta["nf"]) == [4000] and list(meta["nmu"]) == [5]
    assert list(meta["nline"]) == [60] and float(arrays["mu"][-1]) == 1.0
