"""Seeded inputs of the line-rate integrals (bound_bound_rates, one direction): a
few lines of the synthetic atom of dispatch_amd/synth.py on a shared frequency
grid with ragged windows, LTE populations with departures, Planck occupation
numbers and a perturbed intensity.  Spectra are (nf, ncell), cells fastest."""
import numpy as np

from dispatch_amd import synth

H_KT = synth.H_PLANCK / synth.K_BOLTZ


def planck_occupation(nu, T):
    return 1.0 / np.expm1(H_KT * nu / T)


def trapezoid(x):
    w = np.zeros_like(x)
    if len(x) > 1:
        d = np.diff(x)
        w[:-1] += 0.5 * d
        w[1:] += 0.5 * d
    return w


def make_rates_case(ncell=257, nf=240, nline=6, seed=0, flat=False):
    """flat: Bnu constant over frequency per cell, Inu = Bnu and wnu = nu0 * wq, the
    case whose normalised thermal integral has a closed form."""
    rng = np.random.default_rng(seed)
    atom = synth.make_atom(seed=20261204 + seed)
    sel = np.sort(rng.choice(atom.nline, nline, replace=False))
    T = rng.uniform(4500.0, 8000.0, ncell)
    ne = 10 ** rng.uniform(12.0, 15.0, ncell)
    totn = 10 ** rng.uniform(8.0, 10.0, ncell)
    nstar = synth.lte_populations(atom, T, ne, totn)
    b = rng.uniform(0.5, 2.0, nstar.shape)
    lev_i = atom.ti[sel].astype(np.int32)
    lev_j = atom.tj[sel].astype(np.int32)
    nu0 = atom.nu0[sel]
    gi = atom.g[lev_i - 1]
    Aij = atom.Aji[sel] * atom.g[lev_j - 1] / gi
    dopp = rng.uniform(2e-6, 5e-6, ncell)
    inv_df = (1.0 / dopp).astype(np.float32)
    vproj = rng.uniform(-5e-6, 5e-6, ncell).astype(np.float32)
    e_kT = (synth.EV_K / T).astype(np.float32)
    adamp = (10 ** rng.uniform(-3, -0.5, (nline, ncell))).astype(np.float32)
    # ragged windows: widths 1 (a single point) to 90, anywhere in the grid
    q = np.zeros((nline, nf), np.float32)
    wq = np.zeros((nline, nf))
    wnu = np.zeros((nline, nf))
    iblue = np.zeros(nline, np.int32)
    ired = np.zeros(nline, np.int32)
    for l in range(nline):
        width = 1 if l == 1 else int(rng.integers(min(20, nf), min(90, nf) + 1))
        fs = int(rng.integers(0, nf - width + 1))
        iblue[l], ired[l] = fs + 1, fs + width
        qq = np.linspace(-2.5e-5, 2.5e-5, width) if width > 1 else np.zeros(1)
        q[l, fs:fs + width] = qq
        wq[l, fs:fs + width] = trapezoid(qq.astype(np.float32).astype(np.float64)) if width > 1 else 1e-6
        wnu[l, fs:fs + width] = nu0[l] * wq[l, fs:fs + width] * (1.0 if flat else rng.uniform(0.9, 1.1, width))
    lower_off = np.where(rng.uniform(size=nline) < 0.5, 0.0, rng.uniform(0.01, 0.5, nline)).astype(np.float32)
    nu_f = nu0.mean() * (1.0 + 0.02 * (np.arange(nf) / nf - 0.5))
    if flat:
        Bnu = np.broadcast_to(planck_occupation(nu0.mean(), T)[None, :], (nf, ncell))
        Inu = Bnu
    else:
        Bnu = planck_occupation(nu_f[:, None], T[None, :])
        Inu = Bnu * rng.uniform(0.5, 1.5, Bnu.shape)
    Psi = rng.uniform(0.0, 1.0, (nf, ncell))
    beta = rng.uniform(0.3, 1.0, (nf, ncell))
    pop_ratio = rng.uniform(0.5, 2.0, ncell) * 1e-2
    # Bij such that the preconditioning correction is a tenth to a half of the sums
    phi_peak = 0.5641895835 * float(np.median(inv_df))
    nsi_med = np.median(nstar[lev_i - 1], axis=1)
    Bij = 0.3 / (nsi_med * 1e-2 * phi_peak * 0.5)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    return dict(iblue=iblue, ired=ired, lev_i=lev_i, lev_j=lev_j, Aij=Aij, Bij=Bij, lower_off=lower_off,
                q=q, wnu=wnu, wq=wq, g=atom.g.astype(np.float64), b=b, nstar=nstar, pop_ratio=pop_ratio,
                inv_df=inv_df, vproj=vproj, e_kT=e_kT, adamp=adamp, Inu=f32(Inu), Psi=f32(Psi), Bnu=f32(Bnu),
                beta=f32(beta))


def rate_coefficients(d, gij, gji):
    """line_rate_coefficients' rates from the th and dev integrals, FP64:
    R_ij = n_ref th + dev, R_ji = g_i/g_j ((n_ref + 1) th + dev)."""
    li, lj = d["lev_i"] - 1, d["lev_j"] - 1
    g = d["g"]
    beta_ref = d["nstar"][lj] * g[li][:, None] / (d["nstar"][li] * g[lj][:, None])
    n_ref = beta_ref / (1.0 - beta_ref)
    th, dev = gij.astype(np.float64), gji.astype(np.float64)
    Rij = n_ref * th + dev
    Rji = (g[li] / g[lj])[:, None] * ((n_ref + 1.0) * th + dev)
    return Rij, Rji


def level_products(d):
    """n_i, n_j of the rates (b n* times exp(-lower_off e/kT)), FP64."""
    li, lj = d["lev_i"] - 1, d["lev_j"] - 1
    eoff = np.exp(-d["lower_off"][:, None].astype(np.float64) * d["e_kT"][None, :].astype(np.float64))
    ni = d["b"][li] * d["nstar"][li] * eoff
    nj = d["b"][lj] * d["nstar"][lj] * eoff
    return ni, nj


def make_equal_pair_case(ncell=48, nf=160, nline=5, seed=0):
    """Lines that all share one energy gap (2 eV) and one statistical-weight ratio,
    so beta_ref is the same for every line of a cell, and beta(f) = beta_ref,
    Bnu = beta/(1-beta) across the grid.  Overlapping ragged windows, departures
    b of either sign of 1, some lines with a lower-level offset."""
    rng = np.random.default_rng(seed)
    npair = 4
    nk = 2 * npair
    g = np.tile([2.0, 6.0], npair)
    ev = np.repeat(np.linspace(0.0, 1.5, npair), 2) + np.tile([0.0, 2.0], npair)
    T = rng.uniform(4500.0, 8000.0, ncell)
    e_kT = synth.EV_K / T
    totn = 10 ** rng.uniform(8.0, 10.0, ncell)
    w = g[:, None] * np.exp(-ev[:, None] * e_kT[None, :])
    nstar = w * totn / w.sum(0)
    b = rng.uniform(0.4, 2.5, nstar.shape)
    pick = rng.integers(0, npair, nline)
    lev_i = (2 * pick + 1).astype(np.int32)
    lev_j = (2 * pick + 2).astype(np.int32)
    beta_ref = np.exp(-2.0 * e_kT)
    d = dict(lev_i=lev_i, lev_j=lev_j, g=g, b=b, nstar=nstar, e_kT=e_kT.astype(np.float32))
    d["Aij"] = 10 ** rng.uniform(6, 8, nline)
    nu0 = 2.0 * 2.417989242e14
    q = np.zeros((nline, nf), np.float32)
    wq = np.zeros((nline, nf))
    wnu = np.zeros((nline, nf))
    iblue = np.zeros(nline, np.int32)
    ired = np.zeros(nline, np.int32)
    for l in range(nline):
        width = int(rng.integers(min(30, nf), min(120, nf) + 1))
        fs = int(rng.integers(0, nf - width + 1))
        iblue[l], ired[l] = fs + 1, fs + width
        qq = np.linspace(-2.5e-5, 2.5e-5, width) + rng.uniform(-5e-6, 5e-6)
        q[l, fs:fs + width] = qq
        wq[l, fs:fs + width] = trapezoid(qq.astype(np.float32).astype(np.float64))
        wnu[l, fs:fs + width] = nu0 * wq[l, fs:fs + width]
    d.update(iblue=iblue, ired=ired, q=q, wq=wq, wnu=wnu)
    d["lower_off"] = np.where(rng.uniform(size=nline) < 0.4, 0.0, rng.uniform(0.01, 0.3, nline)).astype(np.float32)
    d["inv_df"] = (1.0 / rng.uniform(2e-6, 5e-6, ncell)).astype(np.float32)
    d["vproj"] = rng.uniform(-5e-6, 5e-6, ncell).astype(np.float32)
    d["adamp"] = (10 ** rng.uniform(-3, -0.5, (nline, ncell))).astype(np.float32)
    d["pop_ratio"] = rng.uniform(0.5, 2.0, ncell) * 1e-2
    phi_peak = 0.5641895835 * float(np.median(d["inv_df"]))
    nsi_med = np.median(nstar[lev_i - 1], axis=1)
    d["Bij"] = 0.3 / (nsi_med * 1e-2 * phi_peak)
    beta = np.broadcast_to(beta_ref[None, :], (nf, ncell)).astype(np.float32)
    d["beta"] = np.ascontiguousarray(beta)
    d["Bnu"] = np.ascontiguousarray((beta.astype(np.float64) / (1.0 - beta)).astype(np.float32))
    d["Inu"] = np.ascontiguousarray((d["Bnu"] * rng.uniform(0.3, 2.0, (nf, ncell))).astype(np.float32))
    d["Psi"] = np.zeros((nf, ncell), np.float32)
    return d


def split_solver_outputs(d, opac, Psi):
    """The solver's split outputs consistent with the tracked opacities: total chi =
    xhat - beta ehat, eta = beta ehat, Irem = Inu - Psi eta, escape = 1 - Psi chi."""
    beta = d["beta"].astype(np.float32)
    chi = (opac["xhat"] - opac["ehat"] * beta).astype(np.float32)
    eta = (opac["ehat"] * beta).astype(np.float32)
    Psi = Psi.astype(np.float32)
    return dict(Irem=(d["Inu"] - Psi * eta).astype(np.float32), escape=(1.0 - Psi * chi).astype(np.float32),
                chi=chi, split_opac1=opac["split_opac1"], split_opac2=opac["split_opac2"],
                split_chi1=opac["split_chi1"], split_chi2=opac["split_chi2"], split_id=opac["split_id"],
                split_chi_id=opac["split_chi_id"])
