"""Pins the oracle's formal solvers by answers the mathematics fixes.

- A constant source function with the matching upstream intensity is a fixed
  point of every linear step (E + W + Psi = 1) and of the upwind interpolation
  (Keys' kernel sums to 1), for any opacity, in 3D.
- Linear S(tau) in an atmosphere whose chi500 falls exponentially: the tau500
  step of short_char2 (SRC/rt_solvers_mod.f90:355-370) is then exact and the
  linear step integrates a linear source exactly, so I = a + b*(tau + mu)
  (Eddington-Barbier) at every depth, in 3D along tilted rays.
- A horizontally homogeneous 3D box equals the 1D column solution (long_char2).
- Rolling the box horizontally rolls the answer (periodic indexing), mirroring
  x with the ray mirrors it.
- The FP32 restatement agrees with its FP64 twin to the 1e-5 the sweep must hold.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sweep_cases import DS, gauss_rays, make_case, max_rel, run_oracle


def test_constant_source_is_a_fixed_point():
    case = make_case(nf=6, nx=9, ny=7, nz=8, seed=3)
    case["S"][:] = 0.0372
    case["I"][case["gz"]] = 0.0372
    I, psi, _, _ = run_oracle(case)
    assert np.max(np.abs(I[1:-1] / 0.0372 - 1.0)) < 2e-6


def _exponential_atmosphere(nx, ny, nz, nf, a=0.02, b=0.013, C=3e-6, Hcells=2.5):
    ds = DS[2]
    H = Hcells * ds
    k = np.arange(nz + 2, dtype=np.float64)           # plane 0 is the bottom ghost
    c500 = C * np.exp(-(k - 1) * ds / H)
    tau = c500 * H                                    # vertical optical depth from above
    chi = np.broadcast_to(c500[:, None, None, None], (nz + 2, ny, nx, nf)).astype(np.float32)
    S = np.broadcast_to((a + b * tau)[:, None, None, None], chi.shape).astype(np.float32)
    return c500, tau, chi.copy(), S.copy()


@pytest.mark.parametrize("iray", [0, 3, 6, 9])
def test_eddington_barbier_linear_source(iray):
    nx, ny, nz, nf = 6, 5, 14, 3
    a, b = 0.02, 0.013
    c500, tau, chi, S = _exponential_atmosphere(nx, ny, nz, nf, a, b)
    vec, _ = gauss_rays()
    v = vec[iray]
    mu = v[2]
    I = np.zeros_like(chi)
    I[0] = a + b * (tau[0] + mu)                       # the exact intensity below the box
    chi500 = np.broadcast_to(c500[1:-1, None, None], (nz, ny, nx)).astype(np.float32).copy()
    chi500_up = np.broadcast_to(c500[:-2, None, None], (nz, ny, nx)).astype(np.float32).copy()
    g500 = np.full((ny, nx), c500[0], np.float32)
    Iout, _, _, _ = O.short_char2(chi, S, I, chi500, chi500_up, g500, 1, DS, v)
    exact = a + b * (tau[1:-1] + mu)
    assert max_rel(Iout[1:-1, 0, 0, 0], exact) < 2e-5
    # the FP64 twin is exact to rounding
    I64, _, _, _ = O.short_char2(chi.astype(np.float64), S.astype(np.float64), I.astype(np.float64),
                                 chi500, chi500_up, g500, 1, DS, v, fp64=True)
    assert max_rel(I64[1:-1, 0, 0, 0], exact) < 1e-6


def test_quadratic_solver_linear_source():
    # constant chi: the trapezoid dtau is exact, the parabola of a linear S is the line
    nx, ny, nz, nf = 1, 1, 20, 2
    chi0, a, b = 1.3e-6, 0.01, 0.02
    mu = 0.6
    k = np.arange(nz + 2, dtype=np.float64)
    tau = chi0 * DS[2] * (nz + 1 - k)                 # from the top plane down
    chi = np.full((nz + 2, ny, nx, nf), chi0, np.float32)
    S = np.broadcast_to((a + b * tau)[:, None, None, None], chi.shape).astype(np.float32).copy()
    I = np.zeros_like(chi)
    I[0] = a + b * (tau[0] + mu)
    v = (np.sqrt(1 - mu * mu), 0.0, mu)
    Iq, _ = O.long_char_quadratic(chi, S, I, 1, DS, v)
    exact = a + b * (tau[1:-1] + mu)
    assert max_rel(Iq[1:-1, 0, 0, 0], exact) < 2e-5


def test_homogeneous_box_equals_columns():
    nx, ny, nz, nf = 7, 6, 12, 4
    rng = np.random.default_rng(11)
    z = np.arange(nz + 2)
    c500 = 2e-7 * np.exp(-0.35 * (z - 1))
    prof = (c500[:, None] * 10 ** rng.uniform(0, 2, nf)).astype(np.float32)
    S1 = (0.05 * np.exp(-0.04 * z)[:, None] * (1 + 0.2 * rng.random((nz + 2, nf)))).astype(np.float32)
    chi = np.broadcast_to(prof[:, None, None, :], (nz + 2, ny, nx, nf)).copy()
    S = np.broadcast_to(S1[:, None, None, :], (nz + 2, ny, nx, nf)).copy()
    I = np.zeros_like(chi)
    I[0] = S[0]
    chi500 = np.broadcast_to(c500[1:-1, None, None], (nz, ny, nx)).astype(np.float32).copy()
    cu = np.broadcast_to(c500[:-2, None, None], (nz, ny, nx)).astype(np.float32).copy()
    g500 = np.full((ny, nx), c500[0], np.float32)
    vec, _ = gauss_rays()
    for v in vec[:12]:
        I3, p3, _, _ = O.short_char2(chi, S, I, chi500, cu, g500, 1, DS, v)
        I1, p1, _, _ = O.long_char2(chi[:, :1, :1], S[:, :1, :1], I[:, :1, :1], chi500[:, :1, :1],
                                    g500[:1, :1], 1, DS, v)
        assert max_rel(I3[1:-1], np.broadcast_to(I1[1:-1], I3[1:-1].shape)) < 1e-6
        assert max_rel(p3, np.broadcast_to(p1, p3.shape)) < 1e-6


def test_periodic_roll_and_mirror():
    case = make_case(nf=5, nx=9, ny=8, nz=6, seed=5, vec=(0.45, 0.3, 0.84))
    I0, psi0, _, _ = run_oracle(case)
    rolled = dict(case)
    for k in ("chi", "S", "I"):
        rolled[k] = np.roll(case[k], (2, 3), axis=(1, 2))
    for k in ("chi500", "chi500_up"):
        rolled[k] = np.roll(case[k], (2, 3), axis=(1, 2))
    rolled["g500"] = np.roll(case["g500"], (2, 3), axis=(0, 1))
    I1, psi1, _, _ = run_oracle(rolled)
    assert np.array_equal(np.roll(I0, (2, 3), axis=(1, 2)), I1)
    # mirror x together with the ray's x component
    mir = dict(case)
    for k in ("chi", "S", "I"):
        mir[k] = case[k][:, :, ::-1].copy()
    for k in ("chi500", "chi500_up"):
        mir[k] = case[k][:, :, ::-1].copy()
    mir["g500"] = case["g500"][:, ::-1].copy()
    mir["vec"] = case["vec"] * np.array([-1.0, 1.0, 1.0])
    I2, _, _, _ = run_oracle(mir)
    assert max_rel(I2[:, :, ::-1], I0) < 1e-6


@pytest.mark.parametrize("seed,vec,eta,maser", [
    (1, (0.3, -0.4, 0.866), False, False),
    (2, (-0.5, 0.2, -0.84), False, False),
    (4, (0.2, 0.9, 0.39), True, False),
    (6, (0.1, -0.97, 0.211), True, True),
])
def test_fp32_matches_fp64_twin(seed, vec, eta, maser):
    case = make_case(nf=6, nx=10, ny=9, nz=12, seed=seed, vec=vec, eta=eta, maser=maser)
    I32, p32, _, _ = run_oracle(case)
    I64, p64, _, _ = run_oracle(case, fp64=True)
    floor = 1e-6 * np.max(np.abs(I64))
    assert max_rel(I32[1:-1], I64[1:-1], floor) < 1e-5
    if not maser:
        assert max_rel(p32, p64, 1e-6 * np.max(np.abs(p64))) < 1e-5


def test_split_rates_parts_add_up():
    case = make_case(nf=5, nx=8, ny=7, nz=6, seed=9)
    I, psi, irem, esc = run_oracle(case, split_rates=True)
    # Inu = Irem + Psi*S away from the switched-off points; escape = 1 - Psi_local
    chi = case["chi"][1:-1]
    S = case["S"][1:-1]
    recon = irem + psi * chi * S
    assert max_rel(recon, I[1:-1], 1e-7) < 1e-5
    assert np.all((esc >= 0) & (esc <= 1.0 + 1e-6))
