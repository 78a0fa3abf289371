"""Pins the oracle's short_char_hermite (SRC/rt_solvers_mod.f90:1917-2248):

* exact for a source linear along the ray: constant chi, Pe and T (no chain-rule
  terms), S = B rho with rho linear in height and the ghost intensity on the
  particular solution I = S - dS/dtau, so every upwind interpolation is exact and the
  quadratic step must return I = S - dS/dtau everywhere;
* a constant source in an isothermal medium stays fixed whatever chi does (the
  parabolic weights sum to 1 - e^-dt);
* the slab's first plane is short_char's linear step, bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as O
from hermite_cases import hermite_inputs
from sweep_cases import DS, make_case


def run(case, h, **kw):
    return O.short_char_hermite(case["chi"], case["S"], case["I"], h["beta"], h["pe4"], h["temp"], h["dpe"],
                                h["dlnt"], h["hknu"], case["dir"], case["ds"], case["vec"], **kw)


@pytest.mark.parametrize("vec", [(0.3, -0.4, 0.866), (-0.61, 0.2, -0.5), (0.9, 0.1, 0.42)])
def test_linear_source_is_exact(vec):
    case = make_case(nf=4, nx=8, ny=7, nz=12, seed=3, vec=vec)
    nz2, ny, nx, nf = case["chi"].shape
    h = hermite_inputs(case, isothermal=True)
    h["pe4"][:] = 10.0
    h["dpe"] = [np.zeros_like(h["pe4"])] * 3
    chi0 = np.float32(1e-6)
    case["chi"][:] = chi0
    mu = case["vec"][2]
    zc = (np.arange(nz2, dtype=np.float64) - 1) * DS[2]          # height of the planes, ghosts included
    Bv = h["beta"][0, 0, 0].astype(np.float64) / (1 - h["beta"][0, 0, 0])
    rho = 0.5 + 2e-7 * zc                                        # linear in height
    S = rho[:, None, None, None] * Bv[None, None, None, :] * np.ones((nz2, ny, nx, nf))
    dSdtau = 2e-7 * Bv * mu / float(chi0)
    case["S"] = S.astype(np.float32)
    Iexact = S - dSdtau
    gz = case["gz"]
    case["I"][:] = 0
    case["I"][gz] = Iexact[gz]
    I, psi, _, _ = run(case, h)
    rel = np.abs(I[1:-1] - Iexact[1:-1]) / np.abs(Iexact[1:-1])
    assert rel.max() < 2e-6, rel.max()
    assert np.abs(dSdtau).max() / np.abs(S).max() > 1e-2  # the slope term matters


def test_constant_source_fixed():
    case = make_case(nf=6, nx=9, ny=8, nz=10, seed=4, vec=(0.5, 0.5, 0.707))
    h = hermite_inputs(case, isothermal=True)
    Bv = (h["beta"][0, 0, 0].astype(np.float64) / (1 - h["beta"][0, 0, 0])).astype(np.float32)
    case["S"][:] = Bv
    case["I"][:] = 0
    case["I"][case["gz"]] = Bv
    I, _, _, _ = run(case, h)
    np.testing.assert_allclose(I[1:-1], np.broadcast_to(Bv, I[1:-1].shape), rtol=3e-6)


@pytest.mark.parametrize("vec", [(-0.3, 0.7, -0.64), (0.3, -0.4, 0.866), (-0.61, 0.2, 0.5), (0.45, 0.35, -0.82)])
def test_first_plane_is_short_chars_step(vec):
    """Both sweep directions and all four signs of the upwind offsets: the stencil
    tables of the quadratic solver are short_char's."""
    case = make_case(nf=5, nx=10, ny=9, nz=6, seed=5, vec=vec)
    h = hermite_inputs(case, seed=1)
    I, psi, _, _ = run(case, h)
    Is, psis, _, _ = O.short_char(case["chi"], case["S"], case["I"], case["dir"], case["ds"], case["vec"])
    first = 1 if case["dir"] > 0 else case["nz"]
    np.testing.assert_array_equal(I[first], Is[first])
    np.testing.assert_array_equal(psi[first - 1], psis[first - 1])
    # past the first plane the quadratic step differs from the linear one
    assert np.abs(I[1:-1] - Is[1:-1]).max() > 0
