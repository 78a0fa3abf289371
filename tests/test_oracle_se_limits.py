"""Pins the oracle's update_populations where the limiter and the population floor
bind (SRC/atom_mod.f90:597-642), over many iterations.

A three-level atom whose transitions all carry given rates (the continuum form of
setup_transition_matrix, SRC/atom_mod.f90:740-851, with no line or collision term)
has a rate matrix that does not depend on the populations, so its statistical
equilibrium n_true is the null vector of that matrix (numpy), and the solve of
every iteration returns the correction x = n_true / n.  Then:

- The limiter L read back from one update (the ratio of a limited level's factor to
  an unlimited level's) is the identity on [1/max_update, max_update], continuous
  with unit slope at both ends, increasing, and saturates at e^-1/max_update as
  x -> 0 and 2*max_update as x -> inf (the stated range of the reference's limiter).
- Repeated updates converge to n_true exactly when no floor binds, and conserve
  particles at every iteration.
- With a level whose equilibrium population lies below the floor
  min_pop_frac*totn, the fixed point is fixed by the order of the steps (limit,
  renormalise to totn, then floor): the floored level holds F = min_pop_frac*totn,
  every other level holds norm*n_true with
  norm = totn / (totn - T + F*L(T/F)), T that level's n_true, so the total is
  norm*(totn - T) + F, above totn.  A floor applied before the
  renormalisation, or a limiter of another shape, gives a different fixed point.
"""
import types

import numpy as np
import pytest

from oracle import oracle as O


def _atom3():
    z = np.zeros(0)
    return types.SimpleNamespace(nk=3, ti=np.array([1, 2, 1], np.int32), tj=np.array([2, 3, 3], np.int32),
                                 nline=0, g=np.ones(3), ev=np.zeros(3), Aji=z, pi=np.zeros(0, np.int32),
                                 pj=np.zeros(0, np.int32))


def _rates(r_up, r_dn):
    """float32 rates (nrad, 1) and the matrix they assemble (row = to, column = from)."""
    up = np.array(r_up, np.float32)[:, None]
    dn = np.array(r_dn, np.float32)[:, None]
    M = np.zeros((3, 3))
    for k, (i, j) in enumerate([(0, 1), (1, 2), (0, 2)]):
        M[j, i] += float(up[k, 0])
        M[i, j] += float(dn[k, 0])
    M -= np.diag(M.sum(axis=0))
    return up, dn, M


def _equilibrium(M, totn):
    A = M.copy()
    A[0] = 1.0
    rhs = np.zeros(3)
    rhs[0] = totn
    return np.linalg.solve(A, rhs)


def _update(b, nstar, totn, up, dn, **kw):
    at = _atom3()
    one = np.ones(1)
    return O.update_populations(at, 0.0 * one, 5000.0 * one, np.array([totn]), nstar[:, None], b[:, None],
                                up, dn, None, cswitch=0.0, **kw)


TOTN = 1.0e10
NSTAR = TOTN * np.array([0.3, 0.3, 0.4])


def _limiter_probe(xprobe, max_update):
    """L(xprobe) read from one update: level 3 gets correction xprobe, level 1 and 2
    corrections inside the limiter's identity range."""
    up, dn, M = _rates([5.0, 0.7, 2.0], [3.0, 40.0, 9.0])
    T = _equilibrium(M, TOTN)
    n = np.array([T[0] / 1.3, T[1] / 0.8, T[2] / xprobe])
    b = n / NSTAR
    b2, r = _update(b, NSTAR, TOTN, up, dn, max_update=max_update, min_pop_frac=1e-300)
    assert r["status"] == 0
    g = b2[:, 0] / b
    return 1.3 * g[2] / g[0], (g[0] / 1.3) / (g[1] / 0.8)


@pytest.mark.parametrize("max_update", [100.0, 7.0])
def test_limiter_read_back_from_the_update(max_update):
    eup, elo = max_update, 1.0 / max_update
    L = lambda x: _limiter_probe(x, max_update)[0]
    for x in (elo * 1.0001, 0.3, 1.0, 2.5, eup * 0.9999):
        assert abs(L(x) / x - 1) < 1e-12
    # the two unlimited levels see the same renormalisation
    assert abs(_limiter_probe(0.5, max_update)[1] - 1) < 1e-12
    h = 1e-6
    for j in (elo, eup):
        lo, hi = L(j * (1 - h)), L(j * (1 + h))
        assert abs(lo / j - 1) < 2e-6 and abs(hi / j - 1) < 2e-6                 # continuous
        assert abs((hi - lo) / (2 * h * j) - 1) < 1e-4                           # unit slope
    xs = np.concatenate([np.geomspace(1e-9, elo, 12), np.geomspace(eup, 1e9, 12)])
    vals = np.array([L(x) for x in xs])
    assert np.all(np.diff(vals[:12]) > 0) and np.all(np.diff(vals[12:]) > 0)    # increasing
    assert abs(L(1e-12) / (elo * np.exp(-1.0)) - 1) < 1e-6                        # 0.37 / max_update
    assert abs(L(1e12) / (2 * eup) - 1) < 1e-6                                    # 2 max_update


def test_iteration_converges_to_equilibrium_and_conserves():
    up, dn, M = _rates([5.0, 0.7, 2.0], [3.0, 40.0, 9.0])
    T = _equilibrium(M, TOTN)
    b = np.array([1e-4, 30.0, 3.0])           # level 1 needs a factor > max_update
    for it in range(12):
        b2, r = _update(b, NSTAR, TOTN, up, dn, max_update=100.0, min_pop_frac=1e-30)
        b = b2[:, 0]
        assert abs((b * NSTAR).sum() / TOTN - 1) < 1e-13
        if it == 0:
            assert r["nclamp"] > 0
    assert np.max(np.abs(b * NSTAR / T - 1)) < 1e-12
    assert abs(r["maxerr"]) < 1e-12


@pytest.mark.parametrize("ratio", [0.3, 0.003])
def test_fixed_point_with_the_floor_binding(ratio):
    """level 3's equilibrium population T3 = ratio * F with F the floor population."""
    mpf = 0.1
    F = mpf * TOTN                                # floor population of level 3 (its b floor is 0.25 < 1)
    # rates with n_true,3 = ratio * F: level 3 empties fast into 1 and 2
    T3 = ratio * F
    k_out = 50.0
    k_in = k_out * T3 / (TOTN - T3)               # from 1 and from 2 alike
    up, dn, M = _rates([4.0, k_in, k_in], [6.0, k_out / 2, k_out / 2])
    T = _equilibrium(M, TOTN)
    assert abs(T[2] / T3 - 1) < 1e-5
    max_update = 100.0
    elo = 1 / max_update
    b = np.ones(3)
    for _ in range(200):
        b2, r = _update(b, NSTAR, TOTN, up, dn, max_update=max_update, min_pop_frac=mpf)
        b = b2[:, 0]
    x3 = T[2] / F
    Lx = x3 if x3 >= elo else elo * np.exp(x3 / elo - 1)     # the saturating branch below 1/max_update
    norm = TOTN / (TOTN - T[2] + F * Lx)
    n = b * NSTAR
    assert abs(n[2] / F - 1) < 1e-12
    assert np.max(np.abs(n[:2] / (norm * T[:2]) - 1)) < 1e-10
    assert abs(n.sum() / (norm * (TOTN - T[2]) + F) - 1) < 1e-12
    assert n.sum() > TOTN * (1 + 0.5 * mpf)        # the floor adds particles after the renormalisation
    assert abs(r["maxerr"]) < 1e-10                # and the reported change vanishes there
