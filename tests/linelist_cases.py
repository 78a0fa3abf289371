"""Seeded inputs of the line-list composition (add_linelist_opacities): a few
species of synthetic line lists (two isotopes for one of them, one NLTE species
with departures) over a block of cells on an equidistant or uneven ln(lambda)
grid, with the working xhat/ehat holding a continuum."""
import numpy as np


def lines_of(rng, nl, llam, dl, nlev=0):
    llam0 = np.sort(rng.uniform(llam[0] - 30 * dl, llam[-1] + 30 * dl, nl))
    gpref = (rng.uniform(-3, 0, nl) * np.log(10) - 37.7).astype(np.float32)
    gpref[:2] += 11.0
    return dict(llam0=llam0, lower_ev=rng.uniform(0, 5, nl).astype(np.float32), gpref=gpref,
                grad=(10 ** rng.uniform(-9, -7.5, nl)).astype(np.float32),
                gw1=(10 ** rng.uniform(-30, -28.5, nl)).astype(np.float32),
                gw2=rng.uniform(0.3, 0.4, nl).astype(np.float32),
                gstk=(10 ** rng.uniform(-24, -22, nl) * rng.choice([1, -1], nl)).astype(np.float32),
                unsold=(rng.random(nl) < 0.3).astype(np.int32),
                lev_i=rng.integers(0, nlev + 1, nl).astype(np.int32) if nlev else np.zeros(nl, np.int32),
                lev_j=rng.integers(0, nlev + 1, nl).astype(np.int32) if nlev else np.zeros(nl, np.int32))


def make_case(seed=0, ncell=96, nf=400, equidistant=True, scat=True):
    rng = np.random.default_rng(seed)
    dl = np.log(5010 / 5000) / (nf - 1)
    if equidistant:
        llam = np.log(5000.0) + np.arange(nf) * dl
    else:
        llam = np.log(5000.0) + np.concatenate([[0.0], np.cumsum(dl * rng.uniform(0.3, 1.7, nf - 1))])
    T = rng.uniform(4000, 9000, ncell).astype(np.float32)
    cells = dict(temp=T, vmic=rng.uniform(0, 2, ncell).astype(np.float32),
                 vproj=(rng.uniform(-1, 1, ncell) * 4 * dl).astype(np.float32),
                 ekt=(11604.518 / T).astype(np.float32), ne=(10 ** rng.uniform(12, 15, ncell)).astype(np.float32),
                 tothi=(10 ** rng.uniform(15, 17, ncell)).astype(np.float32))
    cells["spe"] = (cells["ne"] * T ** (1 / 6)).astype(np.float32)
    cells["vmic"][:5] = 0.0
    species = []
    for s, (nl, niso, nlte, group) in enumerate([(40, 1, False, 1), (30, 2, False, 2), (25, 1, True, 1)]):
        nlev = 4 if nlte else 0
        isos = [(lines_of(rng, nl, llam, dl, nlev), rng.uniform(20, 60), 1.0 if k == 0 else rng.uniform(0.05, 0.3))
                for k in range(niso)]
        sp = dict(isotopes=isos, n=10 ** rng.uniform(8, 12, ncell), part=rng.uniform(0.5, 4, ncell),
                  scale=1.0 if s != 1 else 0.7, scat_group=0 if nlte else group)
        sp["n"][:3] = 0.0               # absent: floored to tiny, no opacity
        if nlte:
            sp["dep"] = rng.uniform(0.5, 2.0, (nlev, ncell)).astype(np.float32)
            sp["nlev"] = nlev
        species.append(sp)
    cont = (1e-6 * 10 ** rng.uniform(-0.5, 0.5, (ncell, 1)) * (1 + 0.1 * np.linspace(0, 1, nf))[None, :])
    all_chi = cont.astype(np.float32)
    all_eta = (all_chi * rng.uniform(0.5, 1.0, all_chi.shape)).astype(np.float32)
    all_chi[7, :] = 0.0               # a cell without continuum: logref = -huge
    out = dict(species=species, llam=llam, all_chi=all_chi, all_eta=all_eta, cells=cells, equidistant=equidistant)
    if scat:
        out["one_m_eps"] = rng.uniform(0, 1, (2, ncell, nf)).astype(np.float32)
        out["chi_scat"] = np.zeros((ncell, nf), np.float32)
        out["beta"] = rng.uniform(0.01, 0.5, (ncell, nf)).astype(np.float32)
    return out
