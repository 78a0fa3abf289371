"""Pins the oracle's long_char_hermite (SRC/rt_solvers_mod.f90:1052-1504), the
cubic-Bezier column solver:

* bezier3_weights: each weight against the Bernstein integral
  w_k(t) = t int_0^1 b_k(u) e^{-t u} du by adaptive quadrature in FP64 over
  -3 <= t <= 87 (the reference states 8 FP32 ulp there), and the four summing to
  1 - e^-t;
* a source linear in tau (S = rho B, rho linear in height, chi, Pe, T constant, the
  ghost intensity on the particular solution) gives I = S - dS/dtau on every plane,
  for the three boundary kinds, both directions and nz = 2 (the last plane's
  stencil reaching the ghost) as well as longer columns;
* with S = 0 and chi = q Pe, q quadratic and Pe linear in height with its exact
  gradient, I = I_ghost exp(-tau) with tau the exact integral of the cubic chi (the
  product-rule slopes and the Euler-Maclaurin step are exact for it);
* the Lambda diagonal, the reference's bounded surrogate of dI/dS: 0 <= chi Psi <= 1
  on a random stratified box, and its isothermal constant-chi value
  w1/3 + 2 w2/3 + w3 (w1/3 + w2 + w3 on the first plane below a zeroed top ghost)
  with the weights from quadrature.
"""
import numpy as np
import pytest
from scipy.integrate import quad

from oracle import oracle as O
from hermite_cases import H_K, hermite_inputs
from sweep_cases import DS, make_case

BERN = [lambda u: u ** 3, lambda u: 3 * u * u * (1 - u), lambda u: 3 * u * (1 - u) ** 2, lambda u: (1 - u) ** 3]


def exact_weights(t):
    t = float(t)
    return [t * quad(lambda u: b(u) * np.exp(-t * u), 0.0, 1.0, epsabs=0.0, epsrel=1e-13, limit=200)[0]
            for b in BERN]


def test_bezier_weights_against_quadrature():
    ts = np.concatenate([np.linspace(-3, 3, 61), [2.9999, 3.0, 3.0001, 1e-3, -1e-3, 1e-6],
                         np.geomspace(3.01, 87, 40)]).astype(np.float32)
    E, *w = O.bezier3_weights(ts)
    worst = 0.0
    for i, t in enumerate(ts):
        ex = exact_weights(t)
        for k in range(4):
            ulp = np.spacing(np.float32(abs(ex[k])))
            worst = max(worst, abs(float(w[k][i]) - ex[k]) / float(ulp))
    assert worst <= 8.0, worst
    np.testing.assert_allclose(E, np.exp(-ts.astype(np.float64)), rtol=2e-7)


def test_bezier_weights_sum_to_one_minus_e():
    ts = np.concatenate([np.linspace(-3, 3, 120), np.geomspace(3.01, 87, 60)]).astype(np.float32)
    E, *w = O.bezier3_weights(ts)
    tot = sum(x.astype(np.float64) for x in w)
    m0 = -np.expm1(-ts.astype(np.float64))
    assert np.max(np.abs(tot - m0) / np.abs(m0)) < 2e-6


def column_case(nf, nx, ny, nz, vec, seed=0):
    rng = np.random.default_rng(seed)
    vec = np.asarray(vec, np.float64)
    vec = vec / np.linalg.norm(vec)
    dir = 1 if vec[2] > 0 else -1
    gz = 0 if dir > 0 else nz + 1
    shape = (nz + 2, ny, nx, nf)
    T = 5800.0 + 300.0 * rng.random((ny, nx))
    nu = np.linspace(5.9e14, 6.0e14, nf)
    hknu = (H_K * nu).astype(np.float32)
    beta_c = np.exp(-hknu[None, None, :].astype(np.float64) / T[..., None]).astype(np.float32)
    return dict(vec=vec, dir=dir, gz=gz, shape=shape, T=T, hknu=hknu, beta_c=beta_c, rng=rng)


def ghost_arrays(kind, pe_g, dpe_g, T, ny, nx):
    if kind == "top":
        return {}
    return dict(gpe=np.full((ny, nx), pe_g, np.float32), gdpe=np.full((ny, nx), dpe_g, np.float32),
                gtemp=T.astype(np.float32), gdlnt=np.zeros((ny, nx), np.float32))


@pytest.mark.parametrize("kind", ["top", "bottom", "upstream"])
@pytest.mark.parametrize("vec", [(0.3, -0.4, 0.866), (-0.2, 0.1, -0.5)])
@pytest.mark.parametrize("nz", [2, 7])
def test_linear_source_is_exact(kind, vec, nz):
    nf, nx, ny = 5, 3, 2
    c = column_case(nf, nx, ny, nz, vec, seed=nz)
    nz2 = nz + 2
    mu = c["vec"][2]
    chi0 = np.float32(1.3e-6)
    chi = np.full(c["shape"], chi0, np.float32)
    beta = np.broadcast_to(c["beta_c"], (nz, ny, nx, nf)).astype(np.float32)
    Bv = beta[0].astype(np.float64) / (1 - beta[0].astype(np.float64))          # (ny, nx, nf)
    zc = (np.arange(nz2, dtype=np.float64) - 1) * DS[2]
    rho = 0.5 + 2e-7 * zc
    S = rho[:, None, None, None] * Bv[None]
    dSdtau = 2e-7 * mu * Bv[None] / float(chi0)
    Iex = S - dSdtau
    I0 = np.zeros(c["shape"], np.float32)
    I0[c["gz"]] = Iex[c["gz"]]
    pe = np.full((nz, ny, nx), 12.0, np.float32)
    zero = np.zeros((nz, ny, nx), np.float32)
    temp = np.broadcast_to(c["T"], (nz, ny, nx)).astype(np.float32)
    I, psi = O.long_char_hermite(chi, S.astype(np.float32), I0, beta, pe, temp, zero, zero, c["hknu"], c["dir"], DS,
                                 c["vec"], ghost=kind, **ghost_arrays(kind, 12.0, 0.0, c["T"], ny, nx))
    rel = np.abs(I[1:-1] - Iex[1:-1]) / np.abs(Iex[1:-1])
    assert rel.max() < 3e-6, rel.max()
    assert np.abs(dSdtau).max() / np.abs(S).max() > 1e-2
    # isothermal, constant chi: chi Psi = w1/3 + 2 w2/3 + w3 (first plane under a
    # zeroed top ghost: w1/3 + w2 + w3) at dt = dsup chi
    dt = float(np.float32(np.float32(DS[2]) / abs(mu))) * float(chi0)
    w = exact_weights(dt)
    lam = np.full(nz, w[1] / 3 + 2 * w[2] / 3 + w[3])
    first = 0 if c["dir"] > 0 else nz - 1
    if kind == "top":
        lam[first] = w[1] / 3 + w[2] + w[3]
    got = psi * chi0
    np.testing.assert_allclose(got, np.broadcast_to(lam[:, None, None, None], got.shape), rtol=2e-6)


@pytest.mark.parametrize("kind", ["bottom", "upstream", "top"])
@pytest.mark.parametrize("vec", [(0.5, 0.2, 0.84), (0.1, -0.3, -0.6)])
def test_attenuation_of_a_cubic_opacity(kind, vec):
    nf, nx, ny, nz = 3, 2, 2, 9
    c = column_case(nf, nx, ny, nz, vec, seed=3)
    mu = c["vec"][2]
    dz = DS[2]
    zc = (np.arange(nz + 2, dtype=np.float64) - 1) * dz            # heights of the planes, ghosts included
    H = nz * dz
    # Pe linear, q = chi/Pe quadratic (linear for the top kind, whose first step takes
    # the secant slope of q); chi = q Pe is a cubic in height
    a0 = 1.5e-8
    a1, a2 = -0.5 * a0 / H, (0.0 if kind == "top" else 0.3 * a0 / H ** 2)
    p0, p1 = 20.0, -8.0 / H
    if kind == "top":
        p1 = 0.0
    q = a0 + a1 * zc + a2 * zc ** 2
    pe_all = p0 + p1 * zc
    chi_z = q * pe_all
    chi = np.broadcast_to(chi_z[:, None, None, None], c["shape"]).astype(np.float32)
    S = np.zeros(c["shape"], np.float32)
    I0 = np.zeros(c["shape"], np.float32)
    I0[c["gz"]] = 1.0
    pe = np.broadcast_to(pe_all[1:-1, None, None], (nz, ny, nx)).astype(np.float32)
    dpe = np.full((nz, ny, nx), p1, np.float32)
    zero = np.zeros((nz, ny, nx), np.float32)
    temp = np.broadcast_to(c["T"], (nz, ny, nx)).astype(np.float32)
    beta = np.broadcast_to(c["beta_c"], (nz, ny, nx, nf)).astype(np.float32)
    g = c["gz"]
    I, _ = O.long_char_hermite(chi, S, I0, beta, pe, temp, dpe, zero, c["hknu"], c["dir"], DS, c["vec"], ghost=kind,
                               **ghost_arrays(kind, pe_all[g], p1, c["T"], ny, nx))
    # exact optical depth from the ghost plane, along the ray
    coef = np.polynomial.polynomial.polymul([a0, a1, a2], [p0, p1])
    prim = np.polynomial.polynomial.polyint(coef)
    cum = np.abs(np.polynomial.polynomial.polyval(zc, prim) - np.polynomial.polynomial.polyval(zc[g], prim)) / abs(mu)
    Iex = np.exp(-cum)
    got = I[1:-1, 0, 0, 0].astype(np.float64)
    assert cum[1:-1].max() > 1.0                                    # the column is optically thick
    np.testing.assert_allclose(got, Iex[1:-1], rtol=1e-5)


@pytest.mark.parametrize("kind", ["top", "bottom", "upstream"])
@pytest.mark.parametrize("vec", [(0.3, -0.4, 0.866), (-0.61, 0.2, -0.5)])
def test_lambda_diagonal_is_bounded(kind, vec):
    case = make_case(nf=6, nx=5, ny=4, nz=8, seed=7, vec=vec)
    h = hermite_inputs(case, seed=2)
    ny, nx = h["pe4"].shape[1:]
    gh = {}
    if kind != "top":
        gh = dict(gpe=h["pe4"][0] * 1.1, gdpe=h["dpe"][2][0], gtemp=h["temp"][0] - 50.0, gdlnt=h["dlnt"][2][0])
    I, psi = O.long_char_hermite(case["chi"], case["S"], case["I"], h["beta"], h["pe4"], h["temp"], h["dpe"][2],
                                 h["dlnt"][2], h["hknu"], case["dir"], case["ds"], case["vec"], ghost=kind, **gh)
    chi = case["chi"][1:-1].astype(np.float64)
    lam = psi * chi
    assert np.all(np.isfinite(I)) and np.all(lam >= 0)
    assert np.all(lam <= 1.0)
    assert lam.max() > 0.05


# ---------------------------------------------------------------- long_char_linear

@pytest.mark.parametrize("vec", [(0.0, 0.0, 1.0), (0.0, 0.0, -1.0)])
def test_long_char_linear_is_short_char_on_a_vertical_ray(vec):
    """A vertical ray's 4x4 stencil reads the column itself, so short_char's step is
    long_char_linear's: I, Psi and the split parts agree."""
    case = make_case(nf=5, nx=6, ny=5, nz=7, seed=11, vec=vec, maser=True)
    a = O.short_char(case["chi"], case["S"], case["I"], case["dir"], case["ds"], case["vec"], split_rates=True)
    b = O.long_char_linear(case["chi"], case["S"], case["I"], case["dir"], case["ds"], case["vec"], split_rates=True)
    for x, y in zip(a, b):
        np.testing.assert_allclose(y, x, rtol=1e-6, atol=0)


@pytest.mark.parametrize("vec", [(0.3, -0.4, 0.866), (-0.2, 0.1, -0.5)])
def test_long_char_linear_source_is_exact(vec):
    nf, nx, ny, nz = 4, 3, 2, 9
    c = column_case(nf, nx, ny, nz, vec, seed=5)
    mu = c["vec"][2]
    chi0 = np.float32(8e-7)
    chi = np.full(c["shape"], chi0, np.float32)
    zc = (np.arange(nz + 2, dtype=np.float64) - 1) * DS[2]
    S = np.broadcast_to((0.3 + 1e-7 * zc)[:, None, None, None], c["shape"])
    dSdtau = 1e-7 * mu / float(chi0)
    Iex = S - dSdtau
    I0 = np.zeros(c["shape"], np.float32)
    I0[c["gz"]] = Iex[c["gz"]]
    I, psi, irem, esc = O.long_char_linear(chi, S.astype(np.float32), I0, c["dir"], DS, c["vec"], split_rates=True)
    assert abs(dSdtau) / 0.3 > 1e-2
    np.testing.assert_allclose(I[1:-1], Iex[1:-1], rtol=2e-6)
    # the split parts: I = Irem + (chi Psi) S, escape = 1 - chi Psi
    lam = psi.astype(np.float64) * chi0
    np.testing.assert_allclose(irem + lam * S[1:-1], I[1:-1], rtol=2e-6)
    np.testing.assert_allclose(esc, 1 - lam, rtol=2e-6)


# ---------------------------------------------------------------- long_char_hermite2

@pytest.mark.parametrize("kind", ["top", "bottom", "upstream"])
@pytest.mark.parametrize("vec", [(0.3, -0.4, 0.866), (-0.2, 0.1, -0.5)])
@pytest.mark.parametrize("nz", [2, 3, 7])
def test_hermite2_linear_source_is_exact(kind, vec, nz):
    """The parabolic step with MC-limited (interior), one-sided (last plane, nz >= 3)
    or secant (first plane under the top ghost, last plane at nz = 2) source slopes
    reproduces a source linear in tau on every plane."""
    nf, nx, ny = 4, 3, 2
    c = column_case(nf, nx, ny, nz, vec, seed=nz)
    mu = c["vec"][2]
    chi0 = np.float32(1.1e-6)
    chi = np.full(c["shape"], chi0, np.float32)
    zc = (np.arange(nz + 2, dtype=np.float64) - 1) * DS[2]
    S = np.broadcast_to((0.4 + 1.5e-7 * zc)[:, None, None, None], c["shape"])
    dSdtau = 1.5e-7 * mu / float(chi0)
    Iex = S - dSdtau
    I0 = np.zeros(c["shape"], np.float32)
    I0[c["gz"]] = Iex[c["gz"]]
    pe = np.full((nz, ny, nx), 9.0, np.float32)
    zero = np.zeros((nz, ny, nx), np.float32)
    gh = {} if kind == "top" else dict(gpe=np.full((ny, nx), 9.0, np.float32), gdpe=np.zeros((ny, nx), np.float32))
    I, psi = O.long_char_hermite2(chi, S.astype(np.float32), I0, pe, zero, c["dir"], DS, c["vec"], ghost=kind, **gh)
    assert abs(dSdtau) / 0.4 > 1e-2
    np.testing.assert_allclose(I[1:-1], Iex[1:-1], rtol=2e-6)
    # the Lambda diagonal is the parabolic weight of the target, B / chi, in (0, 1/chi)
    lam = psi * chi0
    assert np.all(lam > 0) and np.all(lam < 1)


@pytest.mark.parametrize("kind", ["bottom", "upstream", "top"])
@pytest.mark.parametrize("vec", [(0.5, 0.2, 0.84), (0.1, -0.3, -0.6)])
def test_hermite2_attenuation_of_a_cubic_opacity(kind, vec):
    nf, nx, ny, nz = 3, 2, 2, 9
    c = column_case(nf, nx, ny, nz, vec, seed=3)
    mu = c["vec"][2]
    dz = DS[2]
    zc = (np.arange(nz + 2, dtype=np.float64) - 1) * dz
    H = nz * dz
    a0 = 1.5e-8
    a1, a2 = -0.5 * a0 / H, (0.0 if kind == "top" else 0.3 * a0 / H ** 2)
    p0, p1 = 20.0, (0.0 if kind == "top" else -8.0 / H)
    q = a0 + a1 * zc + a2 * zc ** 2
    pe_all = p0 + p1 * zc
    chi = np.broadcast_to((q * pe_all)[:, None, None, None], c["shape"]).astype(np.float32)
    S = np.zeros(c["shape"], np.float32)
    I0 = np.zeros(c["shape"], np.float32)
    I0[c["gz"]] = 1.0
    pe = np.broadcast_to(pe_all[1:-1, None, None], (nz, ny, nx)).astype(np.float32)
    dpe = np.full((nz, ny, nx), p1, np.float32)
    g = c["gz"]
    gh = {} if kind == "top" else dict(gpe=np.full((ny, nx), pe_all[g], np.float32),
                                       gdpe=np.full((ny, nx), p1, np.float32))
    I, _ = O.long_char_hermite2(chi, S, I0, pe, dpe, c["dir"], DS, c["vec"], ghost=kind, **gh)
    coef = np.polynomial.polynomial.polymul([a0, a1, a2], [p0, p1])
    prim = np.polynomial.polynomial.polyint(coef)
    cum = np.abs(np.polynomial.polynomial.polyval(zc, prim) - np.polynomial.polynomial.polyval(zc[g], prim)) / abs(mu)
    assert cum[1:-1].max() > 1.0
    np.testing.assert_allclose(I[1:-1, 0, 0, 0].astype(np.float64), np.exp(-cum)[1:-1], rtol=1e-5)
