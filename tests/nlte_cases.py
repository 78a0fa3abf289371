"""Seeded statistical-equilibrium inputs for the oracle and GPU tests: the
synthetic 21-level atom of dispatch_amd/synth.py over a set of cells."""
import numpy as np

from dispatch_amd import synth
from oracle import oracle as O


def make_se_case(ncell=200, seed=0, lte=False, nline=80, ncont=20):
    rng = np.random.default_rng(seed)
    atom = synth.make_atom(nline=nline, ncont=ncont, seed=20261204 + seed)
    T = rng.uniform(4500.0, 8000.0, ncell)
    ne = 10 ** rng.uniform(12.0, 15.0, ncell)
    totn = 10 ** rng.uniform(6.0, 10.0, ncell)
    nstar = synth.lte_populations(atom, T, ne, totn)
    e_kT = synth.EV_K / T
    b = np.ones_like(nstar) if lte else rng.uniform(0.3, 3.0, nstar.shape)
    nrad = len(atom.ti)
    gi = atom.g[atom.ti[:atom.nline] - 1]
    gj = atom.g[atom.tj[:atom.nline] - 1]
    gij = np.zeros((nrad, ncell), np.float32)
    gji = np.zeros((nrad, ncell), np.float32)
    # lines: thermal integral th = A_ji g_j/g_i in radiative equilibrium, deviation dev
    th = (atom.Aji * gj / gi)[:, None] * np.ones(ncell)
    gij[:atom.nline] = th if lte else th * rng.uniform(0.2, 1.5, th.shape)
    gji[:atom.nline] = 0.0 if lte else th * rng.uniform(-0.3, 0.3, th.shape)
    # continua: photoionisation rate and the raw recombination sum (= it, in LTE)
    R = 10 ** rng.uniform(-2, 3, (ncont, ncell))
    gij[atom.nline:] = R
    gji[atom.nline:] = R if lte else R * rng.uniform(0.5, 2.0, R.shape)
    gij, gji = O.normalize_rates(atom, nstar, gij, gji)
    cup = synth.van_regemorter(atom, T, ne, nstar)
    return dict(atom=atom, T=T, ne=ne, totn=totn, nstar=nstar, e_kT=e_kT, b=b, gij=gij, gji=gji, cup=cup)
