"""Pins the oracle's sampled background table (HOP/optab_mod.f90:127-287):

- node slopes, total extinction and its scattering part bit for bit against the
  reference's compiled sample_build_derivatives / sample_interpolate
  (tests/golden/hop_optab_sample.npz, oracle/ref_harness/rt_golden.f90), over
  query cells inside and a little beyond the table and wavelengths at and beyond
  its ends;
- a table linear in wavelength, PC axis 1 and PC axis 2 (any mix, as a sum) is
  reproduced exactly by the tensor-product Hermite with its PCHIP slopes, at any
  query point;
- at the nodes the interpolant returns the node values.
"""
import os

import numpy as np

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hop_optab_sample.npz")


def test_sampled_table_is_the_reference_build():
    g = np.load(GOLD)
    tab = (g["in_lam"], g["in_val"], g["in_scat"], g["in_mu"], g["in_rot"], g["in_c0"], g["in_dc"])
    back, sc, slopes = O.optab_interpolate(*tab, g["in_temp"], g["in_ne"], g["in_chi500"], g["in_lam_q"])
    for name, s in zip(("dl", "dt", "de", "sl", "st", "se"), slopes):
        assert np.array_equal(s, g[name]), name
    assert np.array_equal(back, g["back"])
    assert np.array_equal(sc, g["scat"])


def _pc_query(rng, n, c0, dc, nt, ne, mu, rot):
    p1 = c0[0] + rng.uniform(0, nt - 1, n) * dc[0]
    p2 = c0[1] + rng.uniform(0, ne - 1, n) * dc[1]
    d = rot @ np.vstack([p1, p2])
    return np.exp(d[0] + mu[0]), np.exp(d[1] + mu[1]), p1, p2


def test_linear_table_is_reproduced():
    rng = np.random.default_rng(2)
    nnu, nt, ne = 30, 10, 8
    lam = np.sort(rng.uniform(4000, 6000, nnu))
    mu, c0, dc = np.array([8.7, 31.0]), np.array([-0.9, -2.5]), np.array([0.25, 0.6])
    ang = 0.4
    rot = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    tax = c0[0] + np.arange(nt) * dc[0]
    eax = c0[1] + np.arange(ne) * dc[1]
    a, b, c, e = 2.0, 1e-4, 0.5, -0.3
    val = (a + b * lam[:, None, None] + c * tax[None, :, None] + e * eax[None, None, :]).astype(np.float32)
    scat = (0.1 * val).astype(np.float32)
    temp, nee, p1, p2 = _pc_query(rng, 400, c0, dc, nt, ne, mu, rot)
    lq = np.sort(rng.uniform(lam[0], lam[-1], 50))
    back, sc, _ = O.optab_interpolate(lam, val, scat, mu, rot, c0, dc, temp, nee, np.ones(400, np.float32), lq)
    exact = a + b * lq[None, :] + c * p1[:, None] + e * p2[:, None]
    assert np.max(np.abs(back / (1.1 * exact) - 1)) < 3e-6
    assert np.max(np.abs(sc / (0.1 * exact) - 1)) < 3e-6


def test_nodes_return_node_values():
    g = np.load(GOLD)
    lam, val, mu, rot, c0, dc = g["in_lam"], g["in_val"], g["in_mu"], g["in_rot"], g["in_c0"], g["in_dc"]
    nnu, nt, ne = val.shape
    it, je = np.meshgrid(np.arange(nt), np.arange(ne), indexing="ij")
    p1 = c0[0] + it.ravel() * dc[0]
    p2 = c0[1] + je.ravel() * dc[1]
    d = rot @ np.vstack([p1, p2])
    temp, nee = np.exp(d[0] + mu[0]), np.exp(d[1] + mu[1])
    zero = np.zeros_like(val)
    back, _, _ = O.optab_interpolate(lam, val, zero, mu, rot, c0, dc, temp, nee, np.ones(temp.size, np.float32), lam)
    want = val[:, it.ravel(), je.ravel()].T
    # the PC coordinates of a node go through exp/log and the rotation: FP64 rounding
    assert np.max(np.abs(back / want - 1)) < 1e-5
