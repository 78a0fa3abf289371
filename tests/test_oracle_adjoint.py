"""Pins the oracle's reverse of short_char2 (oracle/m3d_adjoint.c) by central finite
differences of the oracle's FP64 forward (SRC/rt_solvers_mod.f90:224-507): the
derivative of a weighted sum of every plane's intensity with respect to chi and S
of interior cells and of the upstream ghost plane, and to the ghost plane's
intensity, in both sweep directions; and the weight derivatives against differences
of the closed forms E = exp(-t), Psi = 1 - (1 - E)/t, W = 1 - E - Psi."""
import numpy as np
import pytest

from oracle import oracle as O
from sweep_cases import make_case


def closed_weights(t):
    E = np.exp(-t)
    u0 = -np.expm1(-t)
    ps = 1.0 - u0 / t
    return E, u0 - ps, ps


@pytest.mark.parametrize("t", [-1.5, -0.2, 1e-3, 0.03, 0.049, 0.051, 0.4, 2.0, 11.0, 40.0])
def test_weight_derivatives(t):
    E, W, Ps, dE, dW, dPs = O.lin_weights_deriv(t)
    ref = closed_weights(t)
    np.testing.assert_allclose([E, W, Ps], ref, rtol=1e-9, atol=1e-15)
    h = 1e-4 * max(abs(t), 1e-2)
    hi, lo = closed_weights(t + h), closed_weights(t - h)
    fd = [(a - b) / (2 * h) for a, b in zip(hi, lo)]
    np.testing.assert_allclose([dE, dW, dPs], fd, rtol=1e-6, atol=1e-12)


def forward(case, chi, S, I):
    return O.short_char2(chi, S, I, case["chi500"], case["chi500_up"], case["g500"], case["dir"], case["ds"],
                         case["vec"], get_rates=False, fp64=True)[0]


@pytest.mark.parametrize("seed,vec", [(1, (0.3, -0.4, 0.866)), (2, (-0.5, 0.2, -0.84)),
                                      (3, (0.61, 0.05, -0.79)), (4, (-0.2, -0.7, 0.69))])
def test_adjoint_matches_finite_differences(seed, vec):
    case = make_case(nf=3, nx=7, ny=6, nz=5, seed=seed, vec=vec, top_ghost_zero=False)
    chi = case["chi"].astype(np.float64)
    S = case["S"].astype(np.float64)
    I0 = case["I"].astype(np.float64)
    rng = np.random.default_rng(seed)
    w = np.zeros_like(chi)
    w[1:-1] = rng.uniform(-1.0, 1.0, w[1:-1].shape)
    loss = lambda c, s, i: float(np.sum(w * forward(case, c, s, i)))
    I = forward(case, chi, S, I0)
    Ib, cb, sb = O.short_char2_adjoint(chi, S, I, case["chi500"], case["chi500_up"], case["g500"], w,
                                       case["dir"], case["ds"], case["vec"])
    gz = case["gz"]
    nz2, ny, nx, nf = chi.shape
    picks = [(int(rng.integers(1, nz2 - 1)), int(rng.integers(ny)), int(rng.integers(nx)), int(rng.integers(nf)))
             for _ in range(10)]
    picks += [(gz, int(rng.integers(ny)), int(rng.integers(nx)), int(rng.integers(nf))) for _ in range(4)]
    scale = {k: np.abs(v).max() for k, v in (("chi", cb), ("S", sb), ("I", Ib[gz]))}
    for k in picks:
        for name, arr, grad in (("chi", chi, cb), ("S", S, sb), ("I", I0, Ib)):
            if name == "I" and k[0] != gz:
                continue
            h = 1e-6 * abs(arr[k])
            up, dn = arr.copy(), arr.copy()
            up[k] += h
            dn[k] -= h
            args = {"chi": (up, S, I0), "S": (chi, up, I0), "I": (chi, S, up)}[name]
            argd = {"chi": (dn, S, I0), "S": (chi, dn, I0), "I": (chi, S, dn)}[name]
            fd = (loss(*args) - loss(*argd)) / (2 * h)
            assert abs(grad[k] - fd) <= 1e-6 * scale[name] + 1e-5 * abs(fd), (name, k, grad[k], fd)
    # the cotangents reach the ghost plane and every interior plane
    assert np.abs(cb[gz]).max() > 0 and np.abs(sb[gz]).max() > 0 and np.abs(Ib[gz]).max() > 0
