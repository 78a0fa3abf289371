"""Pins the oracle's convert_to_work_units / work_to_transfer_units
(SRC/rt_patch_mod.f90:3518-3536, 3571-3651): the round trip from a physical
(chi_cont, eta) pair through the working units back to (chi, S = eta/chi); the
maser floor (|chi| >= maser_chi_floor |xhat| with the sign kept, the count of floored
points) and its suppress_masers form; the split form equal to the plain one when
the tracked parts sum to ehat; the line-scattering source added to eta.
"""
import numpy as np

from oracle import oracle as O


def fields(n=5000, seed=0):
    rng = np.random.default_rng(seed)
    beta = rng.uniform(0.05, 0.95, n).astype(np.float32)
    cont = (10 ** rng.uniform(-8, -4, n)).astype(np.float32)
    eta = (cont * rng.uniform(0.01, 3.0, n)).astype(np.float32)
    return rng, beta, cont, eta


def test_round_trip():
    rng, beta, cont, eta = fields()
    xhat, ehat = O.convert_to_work_units(cont, eta, beta)
    np.testing.assert_array_equal(xhat, cont + eta)
    np.testing.assert_array_equal(ehat, eta / beta)
    chi, S, nfl = O.work_to_transfer(xhat, ehat, beta)
    assert nfl == 0
    # chi = (cont + eta) - (eta/beta) beta = cont, to the rounding of the three steps
    np.testing.assert_allclose(chi, cont, rtol=0, atol=4 * np.spacing(xhat).max())
    np.testing.assert_allclose(S, (ehat * beta) / chi, rtol=1e-6)


def test_maser_floor_and_suppress():
    rng, beta, cont, eta = fields(seed=1)
    xhat = (cont + eta).astype(np.float32)
    # some points inverted or nearly cancelling: ehat*beta close to or above xhat
    ehat = (xhat / beta * rng.uniform(0.9, 1.1, xhat.size)).astype(np.float32)
    chi, S, nfl = O.work_to_transfer(xhat, ehat, beta, maser_chi_floor=1e-3)
    raw = xhat - ehat * beta
    floor = np.float32(1e-3) * np.abs(xhat)
    low = np.abs(raw) < floor
    assert low.any() and (~low).any()
    assert nfl == int(low.sum())
    np.testing.assert_array_equal(np.abs(chi[low]), floor[low])
    assert np.all(np.sign(chi[low]) == np.where(np.signbit(raw[low]), -1, 1))
    np.testing.assert_array_equal(chi[~low], raw[~low])
    np.testing.assert_allclose(S, (ehat * beta) / chi, rtol=1e-6)
    sup, _, nsup = O.work_to_transfer(xhat, ehat, beta, suppress_masers=True, maser_chi_floor=1e-3)
    assert nsup == 0
    np.testing.assert_array_equal(sup, np.maximum(raw, floor))


def test_split_and_scattering():
    rng, beta, cont, eta = fields(seed=2)
    xhat, ehat = O.convert_to_work_units(cont, eta, beta)
    part = rng.uniform(0.2, 0.8, ehat.size).astype(np.float32)
    so1 = (ehat * part).astype(np.float32)
    so2 = (ehat - so1).astype(np.float32)
    a = O.work_to_transfer(xhat, ehat, beta)
    b = O.work_to_transfer(xhat, ehat, beta, split_opac1=so1, split_opac2=so2)
    np.testing.assert_allclose(b[0], a[0], rtol=1e-5, atol=4 * np.spacing(xhat).max())
    np.testing.assert_allclose(b[1], a[1], rtol=2e-5)
    cs = (cont * rng.uniform(0, 0.5, cont.size)).astype(np.float32)
    dJ = rng.uniform(-1e-2, 1e-2, cont.size).astype(np.float32)
    c = O.work_to_transfer(xhat, ehat, beta, chi_scat=cs, dJnu=dJ)
    np.testing.assert_array_equal(c[0], a[0])  # scattering changes eta only
    np.testing.assert_allclose(c[1] * c[0], a[1] * a[0] + cs * dJ, rtol=1e-5, atol=1e-6 * np.abs(a[1] * a[0]).max())
