"""Pins the oracle's statistical-equilibrium update.

- lineq_no_pivot is bit-identical to the reference's compiled routine
  (tests/golden/hop_lineq.npz) and agrees with numpy's LU solve.
- With J = B_nu (thermal line integrals th = A_ji g_j/g_i, no deviation, recombination
  sums equal to the photoionisation sums) and collisions in detailed balance, LTE
  is the fixed point: every radiative and collisional pair satisfies
  n*_i R_ij = n*_j R_ji (SRC/atom_mod.f90:853-861), so b = 1 stays 1.
- The update conserves particles, sum b n* = totn (SRC/atom_mod.f90:617-631), and
  the solve it uses is numpy's solve of the same assembled matrix.
- The limiter (SRC/atom_mod.f90:597-601) is continuous with unit slope at 1/max_update
  and max_update, and bounds the per-iteration factor by 2*max_update.
"""
import os

import numpy as np

from nlte_cases import make_se_case
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_lineq_no_pivot_is_reference_build():
    g = np.load(os.path.join(GOLD, "hop_lineq.npz"))
    for A, b, x in zip(g["in_A"], g["in_b"], g["x"]):
        assert np.array_equal(O.lineq_no_pivot(A, b), x)


def test_lte_is_a_fixed_point():
    c = make_se_case(ncell=60, seed=1, lte=True)
    b, r = O.update_populations(c["atom"], c["e_kT"], c["T"], c["totn"], c["nstar"], c["b"], c["gij"],
                                c["gji"], c["cup"])
    assert r["status"] == 0
    assert np.max(np.abs(b - 1.0)) < 1e-5
    assert abs(r["maxerr"]) < 1e-5


def test_detailed_balance_of_the_assembled_matrix():
    c = make_se_case(ncell=10, seed=2, lte=True)
    for cell in range(10):
        w, e = O.transition_matrix(c["atom"], cell, c["e_kT"], c["totn"], c["nstar"], c["b"], c["gij"],
                                   c["gji"], c["cup"])
        # every row but the conservation row: the rates out of and into each level balance
        pos = int(np.argmax(e))
        rows = [i for i in range(w.shape[0]) if i != pos]
        resid = w[rows].sum(axis=1)
        scale = np.abs(w[rows]).max(axis=1)
        assert np.max(np.abs(resid) / scale) < 1e-6


def test_update_conserves_particles_and_uses_the_true_solve():
    c = make_se_case(ncell=40, seed=3)
    b, r = O.update_populations(c["atom"], c["e_kT"], c["T"], c["totn"], c["nstar"], c["b"], c["gij"],
                                c["gji"], c["cup"], max_update=1e30)
    assert r["status"] == 0
    assert np.allclose((b * c["nstar"]).sum(axis=0), c["totn"], rtol=1e-12)
    for cell in range(0, 40, 7):
        w, e = O.transition_matrix(c["atom"], cell, c["e_kT"], c["totn"], c["nstar"], c["b"], c["gij"],
                                   c["gji"], c["cup"])
        x = np.linalg.solve(w, e)
        # without the limiter the new b is b*x renormalised (x already conserves)
        assert np.allclose(b[:, cell], c["b"][:, cell] * x, rtol=1e-8)


def test_limiter_bounds_and_measure():
    c = make_se_case(ncell=80, seed=4)
    b0 = c["b"].copy()
    b, r = O.update_populations(c["atom"], c["e_kT"], c["T"], c["totn"], c["nstar"], b0, c["gij"], c["gji"],
                                c["cup"], max_update=3.0)
    assert r["status"] == 0 and r["nclamp"] > 0
    assert np.all(b > 0)
    ratio = b / b0
    # the factor before renormalisation is below 2*max_update; after it the largest
    # relative change is what maxerr reports, at its cell and level
    e = ratio - 1.0
    i, j = np.unravel_index(np.argmax(np.abs(e).T), e.T.shape)
    assert abs(abs(r["maxerr"]) - np.abs(e).max()) < 1e-12 * np.abs(e).max()
    assert (r["epos"], r["elvl"]) == (i, j)
    assert np.allclose(np.abs(r["lvlerr"]), np.abs(e).max(axis=1), rtol=1e-12)


def test_limiter_closed_form():
    eup = 7.0
    elo = 1 / eup
    lim = lambda e: np.where(e < elo, elo * np.exp(np.maximum(e, 0) / elo - 1), np.where(e > eup, eup * (2 - eup / e), e))
    h = 1e-7
    for x in (elo, eup):
        assert abs(lim(np.array(x - h)) - lim(np.array(x + h))) < 3 * h       # continuous, unit slope
        assert abs((lim(np.array(x + h)) - lim(np.array(x - h))) / (2 * h) - 1) < 1e-5
    assert lim(np.array(1e9)) < 2 * eup and lim(np.array(-5.0)) > 0
