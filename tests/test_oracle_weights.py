"""Pins the oracle's scalar pieces against mathematics, not against themselves.

exp_neg / linear_weights / parab_moments are compared with FP64 closed forms to
the ulp bounds the reference states for them (SRC/rt_solvers_mod.f90:772-774 and
857-859: within 4 FP32 ulp over -2 <= t <= 90 / 87).  A dropped or mistyped
polynomial coefficient moves the result by >1e-4 relative and fails here.
cc_kernel is compared with Keys' piecewise cubic-convolution kernel (a = -1/2),
and the Gauss quadrature with the moments it must integrate exactly.
"""
import numpy as np
import pytest

from oracle import oracle as O


def ulp_err(a, exact):
    a = np.asarray(a, np.float64)
    sp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    return np.abs(a - exact) / sp


def psi_exact(t):
    t = np.asarray(t, np.float64)
    u0 = -np.expm1(-t)
    with np.errstate(all="ignore"):
        big = 1.0 - u0 / t
    small = t * (0.5 + t * (-1 / 6 + t * (1 / 24 + t * (-1 / 120 + t * (1 / 720 + t * (-1 / 5040))))))
    return np.where(np.abs(t) < 1e-2, small, big)


T = np.linspace(-2.0, 90.0, 400001).astype(np.float32)


def test_exp_neg_against_exp():
    t = T[T <= 87.0]
    E, u0 = O.exp_neg(t)
    td = t.astype(np.float64)
    assert ulp_err(E, np.exp(-td)).max() <= 2.0
    m = np.abs(td) > 0
    assert ulp_err(u0[m], -np.expm1(-td[m])).max() <= 3.0


def test_exp_neg_clamps_at_87():
    E, u0 = O.exp_neg(np.array([87.0, 100.0, 1e6], np.float32))
    assert np.all(E == E[0]) and np.all(u0 == 1.0)


def test_linear_weights_against_closed_form():
    E, W, P, Es = O.linear_weights(T)
    td = T.astype(np.float64)
    Ee = np.exp(-np.minimum(td, 87.0))
    Pe = psi_exact(td)
    u0 = -np.expm1(-td)
    We = u0 - Pe
    Ese = 1.0 - Pe
    nz = td != 0
    assert ulp_err(E, Ee).max() <= 2.0
    assert ulp_err(P[nz], Pe[nz]).max() <= 4.0
    assert ulp_err(Es, Ese).max() <= 2.0
    # W = u0 - Psi: the reference states 4 ulp; separately rounded FP32 (no fused
    # multiply-add) reaches 4.8 ulp near t = -0.34
    assert ulp_err(W[nz], We[nz]).max() <= 6.0
    # t = 0 exactly: no division by zero leaks out of the thin branch
    E0, W0, P0, Es0 = O.linear_weights(np.zeros(1, np.float32))
    assert E0[0] == 1.0 and W0[0] == 0.0 and P0[0] == 0.0 and Es0[0] == 1.0


def test_linear_step_identities():
    # E + W + Psi = 1 (a constant source passes unchanged), Esc = 1 - Psi
    E, W, P, Es = O.linear_weights(T[T <= 30])
    assert np.max(np.abs(E.astype(np.float64) + W + P - 1.0)) < 4e-7
    assert np.max(np.abs(Es.astype(np.float64) - (1.0 - P.astype(np.float64)))) < 2e-7


def test_parab_moments_against_closed_form():
    t = T[T <= 87.0]
    E, U0, U1, q = O.parab_moments(t)
    td = t.astype(np.float64)
    u0 = -np.expm1(-td)
    u1 = td * psi_exact(td)           # U1 = t - U0 = t*Psi
    with np.errstate(all="ignore"):
        qe = np.where(np.abs(td) < 1e-2,
                      td * td * (1 / 3 + td * (-1 / 12 + td * (1 / 60 + td * (-1 / 360)))),
                      td - 2.0 * u1 / td)
    nz = td != 0
    assert ulp_err(E, np.exp(-td)).max() <= 2.0
    assert ulp_err(U0[nz], u0[nz]).max() <= 3.0
    assert ulp_err(U1[nz], u1[nz]).max() <= 4.0
    assert ulp_err(q[nz], qe[nz]).max() <= 5.0


def keys(d, a=-0.5):
    d = abs(d)
    if d <= 1:
        return (a + 2) * d ** 3 - (a + 3) * d ** 2 + 1
    if d < 2:
        return a * d ** 3 - 5 * a * d ** 2 + 8 * a * d - 4 * a
    return 0.0


@pytest.mark.parametrize("s", [0.0, 0.1, 0.25, 0.5, 0.731, 0.999, 1.0])
def test_cc_kernel_is_keys(s):
    k = O.cc_kernel(s)
    ref = [keys(1 + s), keys(s), keys(1 - s), keys(2 - s)]
    assert np.allclose(k, ref, rtol=0, atol=1e-14)
    assert abs(k.sum() - 1.0) < 1e-15 * 4                   # partition of unity
    assert abs(np.dot(k, [-1, 0, 1, 2]) - s) < 1e-15        # reproduces linear functions


def test_set_stencil_locates_upwind_point():
    ds = (50e5, 40e5, 10e5)
    for vec in ([0.3, -0.2, 0.9327379], [-0.6, 0.5, 0.6244998], [0.0, 0.0, -1.0]):
        vec = np.array(vec) / np.linalg.norm(vec)
        r = O.set_stencil(ds, vec)
        assert r["dsup"] == pytest.approx(ds[2] / abs(vec[2]))
        for kern, upo, c in ((r["xkern"], r["upo1"], 0), (r["ykern"], r["upo2"], 1)):
            disp = -vec[c] / ds[c] * r["dsup"]          # upwind point, in cells
            pts = np.arange(4) + upo                     # offsets of the 4 points
            assert np.dot(kern, pts) == pytest.approx(disp, abs=1e-12)


def test_gauss_quadrature_moments():
    vec, w = O.gauss_quadrature(2, 6)
    assert vec.shape == (24, 3)
    assert w.sum() == pytest.approx(1.0, abs=1e-7)
    assert np.allclose(np.linalg.norm(vec, axis=1), 1.0, atol=1e-7)
    mu = vec[:, 2]
    assert np.dot(w, mu ** 2) == pytest.approx(1 / 3, abs=1e-7)   # 2-point Gauss is exact for mu^2
    assert np.dot(w, mu) == pytest.approx(0.0, abs=1e-12)          # mirrored set
    assert np.dot(w, vec[:, 0] ** 2) == pytest.approx(1 / 3, abs=1e-7)
    # the tabulated default-real nodes of the reference
    assert mu[0] == np.float64(np.float32(0.78867513))
    fvec, fw = O.gauss_quadrature(2, 6, jnu_weights=False)
    assert fvec.shape == (12, 3) and np.all(fvec[:, 2] > 0)
    assert fw.sum() == pytest.approx(0.5, abs=1e-7)                 # int_0^1 mu dmu
