"""Pins the oracle's short_char2 eta step, its split outputs and short_char's
tilted step by answers the mathematics fixes (SRC/rt_solvers_mod.f90:19-189,
409-491).

- The eta step integrates e = x*S = eta/chi500 linearly in tau and divides by the
  mean x of the step (rxb = dt500/dt).  Where x = chi/chi500 is the same at both
  ends of the step (chi proportional to chi500), that is the S step exactly: a
  flagged and an unflagged sweep give the same I, Psi, Irem and escape.
- As chi -> 0 the step becomes the trapezoid rule for eta along the path,
  I = I_up + (eta_up + eta)/2 * ds, exact for an emissivity linear along the ray.
- The rates' outputs are derivatives and parts of I: Psi = dI/deta at fixed chi
  (eta = chi*S), Irem = I - eta*Psi, escape = 1 - chi*Psi, in flagged and
  unflagged cells alike, and Psi = 0, Irem = I, escape = 1 where the step is
  switched off (chi <= 0).
- short_char on tilted rays: with constant chi its step length is exact, the 4x4
  cubic of a horizontally uniform field is exact, and the linear step integrates
  a linear source exactly, so I = a + b*(tau + mu) at every depth (Eddington-
  Barbier); a constant source with the matching upstream intensity stays fixed.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sweep_cases import DS, gauss_rays, make_case, max_rel


def _all_flagged(nz, ny, nx, nf):
    ncell = nz * ny * nx
    row = np.arange(1, ncell + 1, dtype=np.int32).reshape(nz, ny, nx)
    mask = np.ones((ncell, nf), np.uint8)
    return row, mask


def _proportional_case(seed, vec, nf=5, nx=9, ny=8, nz=7):
    """chi = kappa_f * chi500 with chi500 varying in space; S random."""
    case = make_case(nf=nf, nx=nx, ny=ny, nz=nz, seed=seed, vec=vec, top_ghost_zero=False)
    rng = np.random.default_rng(seed + 100)
    kappa = 10 ** (rng.choice([-1.0, 1.0], nf) * rng.uniform(0.4, 1.5, nf))
    nzp = nz + 2
    c500 = np.empty((nzp, ny, nx), np.float64)
    c500[1:-1] = case["chi500"]
    gz = case["gz"]
    c500[gz] = case["g500"]
    other = nz + 1 if gz == 0 else 0
    c500[other] = c500[1] if other == 0 else c500[nz]
    # float64: the FP64 sweep then sees x = kappa to rounding (the FP32 sweep casts)
    case["chi"] = c500[..., None].astype(np.float32).astype(np.float64) * kappa
    # chi500 at the upwind point: the upwind plane's own value, so x is kappa at both ends
    return case, kappa


@pytest.mark.parametrize("seed,vec", [(1, (0.3, -0.4, 0.866)), (2, (-0.45, 0.6, -0.66)),
                                      (3, (0.9, 0.1, 0.42))])
@pytest.mark.parametrize("fp64", [False, True])
def test_eta_step_equals_s_step_when_x_is_constant(seed, vec, fp64):
    case, _ = _proportional_case(seed, vec)
    nz, ny, nx, nf = case["nz"], case["ny"], case["nx"], case["nf"]
    row, mask = _all_flagged(nz, ny, nx, nf)
    args = (case["chi"], case["S"], case["I"], case["chi500"], case["chi500_up"], case["g500"],
            case["dir"], case["ds"], case["vec"])
    I0, p0, r0, e0 = O.short_char2(*args, split_rates=True, fp64=fp64)
    I1, p1, r1, e1 = O.short_char2(*args, eta_row=row, eta_mask=mask, split_rates=True, fp64=fp64)
    tol = 1e-11 if fp64 else 2e-6
    assert max_rel(I1[1:-1], I0[1:-1]) < tol
    assert max_rel(p1, p0) < tol
    assert max_rel(r1, r0, 1e-9 * np.abs(r0).max()) < tol
    assert max_rel(e1, e0) < tol
    # and a planted inversion of rxb (xbar instead of 1/xbar) would not agree
    assert np.all(np.abs(np.log10(case["chi"][1:-1] / case["chi500"][..., None])) > 0.35)


@pytest.mark.parametrize("iray", [0, 4, 13, 21])
@pytest.mark.parametrize("fp64", [False, True])
def test_eta_step_thin_limit_is_trapezoid_of_eta(iray, fp64):
    nx, ny, nz, nf = 6, 5, 9, 3
    vec, _ = gauss_rays()
    v = vec[iray]
    mu = abs(v[2])
    dir = 1 if v[2] > 0 else -1
    c0 = 1.0e-7
    eps = 1.0e-10                    # chi / chi500: the step is 1e-11 thick
    eta0 = np.array([2.0e-9, 7.0e-9, 1.3e-8])
    g = np.array([3.0e-10, -4.0e-10, 1.1e-9])
    k = np.arange(nz + 2, dtype=np.float64)
    pos = k if dir > 0 else (nz + 1 - k)          # planes crossed from the ghost plane
    eta = eta0[None, :] + g[None, :] * pos[:, None]            # (nz+2, nf), linear along the ray
    chi = np.full((nz + 2, ny, nx, nf), eps * c0)
    S = np.broadcast_to((eta / (eps * c0))[:, None, None, :], chi.shape).copy()
    gz = 0 if dir > 0 else nz + 1
    I = np.zeros_like(chi)
    I0 = np.array([0.01, 0.02, 0.03])
    I[gz] = I0
    chi500 = np.full((nz, ny, nx), c0)
    g500 = np.full((ny, nx), c0)
    row, mask = _all_flagged(nz, ny, nx, nf)
    Iout, psi, irem, esc = O.short_char2(chi, S, I, chi500, chi500, g500, dir, DS, v, eta_row=row,
                                         eta_mask=mask, split_rates=True, fp64=fp64)
    dsup = DS[2] / mu
    n = pos[1:-1]                                  # steps taken to reach each plane
    exact = I0[None, :] + dsup * (eta0[None, :] * n[:, None] + g[None, :] * n[:, None] ** 2 / 2)
    got = Iout[1:-1, 2, 3, :]
    assert max_rel(got, exact) < (1e-8 if fp64 else 3e-6)
    # Psi -> dI/deta = ds/2 of the cell's own end of the trapezoid
    assert max_rel(psi[:, 2, 3, :], np.full((nz, nf), 0.5 * dsup)) < (1e-8 if fp64 else 3e-6)


@pytest.mark.parametrize("seed,vec,maser", [(4, (0.2, 0.9, 0.39), False), (6, (0.1, -0.97, 0.211), True),
                                            (7, (-0.6, -0.3, 0.74), True)])
def test_split_outputs_are_derivatives_and_parts_of_I(seed, vec, maser):
    nf, nx, ny, nz = 4, 7, 6, 5
    case = make_case(nf=nf, nx=nx, ny=ny, nz=nz, seed=seed, vec=vec, eta=True, maser=maser,
                     top_ghost_zero=False)
    case["maser"] = maser
    args = lambda S: (case["chi"], S, case["I"], case["chi500"], case["chi500_up"], case["g500"],
                      case["dir"], case["ds"], case["vec"])
    kw = dict(eta_row=case["eta_row"], eta_mask=case["eta_mask"], split_rates=True, fp64=True)
    I, psi, irem, esc = O.short_char2(*args(case["S"]), **kw)
    chi = case["chi"][1:-1].astype(np.float64)
    S = case["S"][1:-1].astype(np.float64)
    Iin = I[1:-1]
    on = chi > 0
    # parts of I
    assert max_rel(irem, Iin - chi * S * psi, 1e-9 * np.abs(Iin).max()) < 1e-9
    assert max_rel(esc[on], 1.0 - chi[on] * psi[on]) < 1e-9
    off = ~on
    assert np.all(psi[off] == 0) and np.all(esc[off] == 1.0) and np.array_equal(irem[off], Iin[off])
    # Psi = dI/deta at fixed chi: raise S of one cell by a part in 1e6 (one cell per
    # run: the cell's change reaches the downstream planes' intensities)
    rng = np.random.default_rng(seed)
    flagged = 0
    unflagged = 0
    for z in range(nz):
        want = z % 2 == 0                # alternate flagged and unflagged cells
        ys, xs = np.nonzero((case["eta_row"][z] > 0) == want)
        k = rng.integers(len(ys))
        y, x = ys[k], xs[k]
        S2 = case["S"].astype(np.float64).copy()
        S2[z + 1, y, x, :] *= 1.0 + 1e-6
        I2, _, _, _ = O.short_char2(*args(S2), **kw)
        for f in range(nf):
            c = chi[z, y, x, f]
            if psi[z, y, x, f] == 0:     # the step is switched off there
                assert c <= 0 or case["maser"]
                continue
            deta = c * (S2[z + 1, y, x, f] - S[z, y, x, f])
            dI = I2[z + 1, y, x, f] - I[z + 1, y, x, f]
            assert abs(dI / deta - psi[z, y, x, f]) <= 1e-6 * abs(psi[z, y, x, f])
            r = case["eta_row"][z, y, x]
            flagged += int(r > 0 and case["eta_mask"][r - 1, f])
            unflagged += int(r == 0)
    assert flagged > 0 and unflagged > 0      # both kinds of cell were probed


@pytest.mark.parametrize("iray", [0, 2, 5, 8, 11])
def test_short_char_tilted_eddington_barbier(iray):
    nx, ny, nz, nf = 6, 5, 16, 2
    chi0, a, b = 1.6e-6, 0.02, 0.013
    vec, _ = gauss_rays()
    v = vec[iray]
    assert v[2] > 0
    mu = v[2]
    k = np.arange(nz + 2, dtype=np.float64)
    tau = chi0 * DS[2] * (nz + 1 - k)            # vertical optical depth from the top plane
    chi = np.full((nz + 2, ny, nx, nf), chi0, np.float32)
    S = np.broadcast_to((a + b * tau)[:, None, None, None], chi.shape).astype(np.float32).copy()
    I = np.zeros_like(chi)
    I[0] = a + b * (tau[0] + mu)                 # the semi-infinite solution below the box
    Iout, psi, _, _ = O.short_char(chi, S, I, 1, DS, v)
    exact = a + b * (tau[1:-1] + mu)
    assert max_rel(Iout[1:-1], np.broadcast_to(exact[:, None, None, None], Iout[1:-1].shape)) < 2e-6
    # Psi of the linear step: 1 - (1 - e^-t)/t over chi, t = chi ds / mu
    t = chi0 * DS[2] / mu
    psi_exact = (1 - (1 - np.exp(-t)) / t) / chi0
    assert max_rel(psi, np.full(psi.shape, psi_exact)) < 2e-6


@pytest.mark.parametrize("seed,vec", [(2, (0.35, -0.2, 0.91)), (5, (-0.8, 0.5, -0.33))])
def test_short_char_tilted_constant_source_fixed(seed, vec):
    case = make_case(nf=4, nx=8, ny=7, nz=6, seed=seed, vec=vec, top_ghost_zero=False)
    S0 = 0.0415
    case["S"][:] = S0
    case["I"][case["gz"]] = S0
    I, _, _, _ = O.short_char(case["chi"], case["S"], case["I"], case["dir"], case["ds"], case["vec"])
    assert np.max(np.abs(I[1:-1] / S0 - 1.0)) < 2e-6
