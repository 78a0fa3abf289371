"""Batched synthesis of uneven plane-parallel LTE atmospheres (m3drt_b1d_*).

The workload: 8192 plane-parallel atmospheres with Teff uniform in 3500-7500 K, log g
uniform in 1-5, depth counts integer-uniform in 48-256 (log tau500 uniform from -5
to +2), frequency counts log-uniform in 200-6000 over 500-510 nm, 3-7 upward mu
(Gauss-Legendre on (0, 1]) and 20-200 lines each, drawn as the single-atmosphere
synthesis draws its 60: centres uniform in the band, log gf uniform in [-3, 0],
excitation energy uniform in 0-5 eV, Voigt damping log-uniform in [1e-3, 1e-1],
microturbulence 1 km/s.  T(tau) = Teff (0.75 (tau + 2/3))^(1/4); gas pressure from
hydrostatic equilibrium with kappa500 proportional to P^(1/2); electron pressure from
a single-ion Saha balance of the electron donors (7.6 eV); the continuum is grey plus
a lambda^3 term.  Line strengths (relative to chi500) are synthetic: gf, the
Boltzmann factor of the lower level and the neutral fraction of the same ion,
scaled so that the strongest lines saturate.  S = B_nu(T).  Each line enters only
inside its window at each depth: where s K(a, v) stays above tol = 1e-4 of the
continuum (the Gaussian core and the Lorentz wing bounds), as add_opacity_equi's
tolerance windows it (HOP/linelist_mod.f90:853-906); weaker lines are skipped.

The reference runs such an atmosphere as one dispatch.x 1D run (long_char2 per
frequency window and direction, SRC/rt_solvers_mod.f90:527-663, with add_opacity_equi
per species); here the whole batch is one persistent launch (m3drt_b1d_plan_*),
against the plain path of one launch per atmosphere (m3drt_b1d_synth_plain).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import m3drt

C_LIGHT = 2.99792458e10
K_BOLTZ = 1.380649e-16
H_PLANCK = 6.62607015e-27
AMU = 1.66053907e-24
EV_K = 11604.518  # K per eV
LAM_REF = 500e-7
TOL = 1e-4


@dataclass
class B1dConfig:
    natm: int = 8192
    seed: int = 20261205
    teff: tuple = (3500.0, 7500.0)
    logg: tuple = (1.0, 5.0)
    nd: tuple = (48, 256)
    nf: tuple = (200, 6000)
    nmu: tuple = (3, 7)
    nline: tuple = (20, 200)
    band: tuple = (500e-7, 510e-7)
    vmic: float = 1e5
    mass_amu: float = 56.0
    chi_ion: float = 7.6
    log_s0: float = 4.0
    mu_rule: str = "gauss"   # or "radau": Gauss-Radau with mu = 1 included


# The single-atmosphere synthesis: Teff 5800 K, log g 4.4, 72 depth points, 4000
# wavelengths over 500-510 nm, 5 Gauss-Radau mu (mu = 1 included), 60 lines.
CONFIG0 = B1dConfig(natm=1, seed=20261202, teff=(5800.0, 5800.0), logg=(4.4, 4.4), nd=(72, 72),
                    nf=(4000, 4000), nmu=(5, 5), nline=(60, 60), mu_rule="radau")


def gauss_mu(n):
    """Gauss-Legendre nodes and weights on (0, 1] (weights sum to 1)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def radau_mu(n):
    """Gauss-Radau nodes and weights on [0, 1] with the node mu = 1 (weights sum to 1;
    exact for polynomials of degree 2n - 2), as radau%init places them
    (solvers/rt/radau_mod.f90:15-115): the free nodes are the roots of
    (P_(n-1) - P_n)/(1 - x) on [-1, 1]."""
    from numpy.polynomial import legendre as Lg
    if n == 1:
        return np.array([1.0]), np.array([1.0])
    c = np.zeros(n + 1)
    c[n - 1], c[n] = 1.0, -1.0
    r = np.sort(np.real(Lg.legroots(c)))
    r = r[np.abs(r - 1.0) > 1e-12]
    x = np.append(r, 1.0)
    pn1 = Lg.legval(x, np.eye(n)[n - 1])
    w = np.where(np.abs(x - 1.0) < 1e-12, 2.0 / n ** 2, (1.0 + x) / (n * pn1) ** 2)
    return 0.5 * (x + 1.0), 0.5 * w


def _offsets(counts):
    off = np.zeros(len(counts), np.int64)
    off[1:] = np.cumsum(counts)[:-1]
    return off


def make_batch(cfg: B1dConfig = B1dConfig(), device="cuda"):
    """The batch's arrays as a dict of contiguous tensors on device (the dtypes of
    m3drt.B1D_DTYPES), with I_top and flux allocated; plus host-side metadata."""
    rng = np.random.default_rng(cfg.seed)
    n = cfg.natm
    teff = rng.uniform(*cfg.teff, n)
    logg = rng.uniform(*cfg.logg, n)
    nd = rng.integers(cfg.nd[0], cfg.nd[1] + 1, n).astype(np.int32)
    nf = np.rint(np.exp(rng.uniform(math.log(cfg.nf[0]), math.log(cfg.nf[1]), n))).astype(np.int32)
    nmu = rng.integers(cfg.nmu[0], cfg.nmu[1] + 1, n).astype(np.int32)
    nline = rng.integers(cfg.nline[0], cfg.nline[1] + 1, n).astype(np.int32)
    d_off, l_off, mu_off, f_off = _offsets(nd), _offsets(nline), _offsets(nmu), _offsets(nf)
    i_off = _offsets(nmu.astype(np.int64) * nf)
    ld_off = _offsets(nline.astype(np.int64) * nd)
    lam0 = np.full(n, cfg.band[0])
    dlam = (cfg.band[1] - cfg.band[0]) / (nf - 1.0)

    # ---- depth points (flat), on the device
    dev = torch.device(device)
    f64 = dict(dtype=torch.float64, device=dev)
    ndt = torch.as_tensor(nd.astype(np.int64), device=dev)
    atm_d = torch.repeat_interleave(torch.arange(n, device=dev), ndt)
    k = torch.arange(int(nd.sum()), device=dev) - torch.as_tensor(d_off, device=dev)[atm_d]
    ndd = ndt[atm_d].double()
    ltau = -5.0 + 7.0 * k.double() / (ndd - 1.0)
    tau = 10.0 ** ltau
    Teff = torch.as_tensor(teff, **f64)[atm_d]
    g = 10.0 ** torch.as_tensor(logg, **f64)[atm_d]
    T = Teff * (0.75 * (tau + 2.0 / 3.0)) ** 0.25
    # hydrostatic with kappa500 = k1 P^(1/2): P^(3/2) = 1.5 g tau / k1 (P ~ 1e5 at tau 1, log g 4.4)
    Pg = (1.5 * g * tau / 1.19e-3) ** (2.0 / 3.0)
    # single-ion Saha of the donors: Pe^2 + Phi Pe - A Pg Phi = 0
    phi = 0.6665 * T ** 2.5 * 10.0 ** (-5040.0 * cfg.chi_ion / T)
    Pe = 0.5 * (-phi + torch.sqrt(phi * phi + 4.0 * 1e-4 * Pg * phi))
    neutral = Pe / (Pe + phi)
    dtau = torch.zeros_like(tau)
    nxt = torch.roll(tau, -1)
    last = k == (ndt[atm_d] - 1)
    dtau = torch.where(last, torch.zeros_like(tau), nxt - tau)
    hckT = H_PLANCK * C_LIGHT / (K_BOLTZ * T)
    vth = torch.sqrt(2.0 * K_BOLTZ * T / (cfg.mass_amu * AMU) + cfg.vmic ** 2)
    cont_a = torch.full_like(T, 0.7)
    cont_b = torch.full_like(T, 0.3)

    # ---- lines (flat)
    nl_tot = int(nline.sum())
    atm_l = np.repeat(np.arange(n), nline)
    lam_l = rng.uniform(cfg.band[0], cfg.band[1], nl_tot)
    loggf = rng.uniform(-3.0, 0.0, nl_tot)
    elow = rng.uniform(0.0, 5.0, nl_tot)
    adamp = np.exp(rng.uniform(math.log(1e-3), math.log(1e-1), nl_tot))
    line_off = (lam_l - lam0[atm_l]).astype(np.float32)
    line_cl = (C_LIGHT / lam_l).astype(np.float32)

    # ---- (line, depth) records, each atmosphere depth-major
    nrec = int((nline.astype(np.int64) * nd).sum())
    cnt = torch.as_tensor(nline.astype(np.int64) * nd, device=dev)
    atm_r = torch.repeat_interleave(torch.arange(n, device=dev), cnt)
    i = torch.arange(nrec, device=dev) - torch.as_tensor(ld_off, device=dev)[atm_r]
    nl_r = torch.as_tensor(nline.astype(np.int64), device=dev)[atm_r]
    d_r = i // nl_r
    l_r = i - d_r * nl_r
    dg = torch.as_tensor(d_off, device=dev)[atm_r] + d_r        # flat depth index
    lg = torch.as_tensor(l_off, device=dev)[atm_r] + l_r        # flat line index
    Tr = T[dg]
    s = 10.0 ** (cfg.log_s0 + torch.as_tensor(loggf, **f64)[lg] - 5040.0 * torch.as_tensor(elow, **f64)[lg] / Tr) \
        * neutral[dg] * torch.sqrt(10.0 / Pe[dg])
    a = torch.as_tensor(adamp, **f64)[lg]
    # window half-width: s K < tol * (the band's smallest continuum) beyond it, with the
    # Gaussian core K = exp(-v^2) and the Lorentz wing K = a / (sqrt(pi) v^2)
    cmin = 0.7 + 0.3 * (cfg.band[0] / LAM_REF) ** 3
    ratio = s / (TOL * cmin)
    vg = torch.sqrt(torch.clamp(torch.log(ratio), min=0.0))
    vl = torch.sqrt(ratio * a / math.sqrt(math.pi))
    lam_lr = torch.as_tensor(lam_l, **f64)[lg]
    reach = torch.maximum(vg, vl) * lam_lr * vth[dg] / C_LIGHT
    reach = torch.where(ratio > 1.0, reach, torch.full_like(reach, -1.0))
    rec = torch.stack([a, s, reach, torch.zeros_like(a)], dim=1).float().contiguous()
    del atm_r, i, nl_r, d_r, l_r, dg, lg, Tr, s, a, ratio, vg, vl, lam_lr, reach

    mus, wfl = [], []
    for m in range(n):
        x, w = (radau_mu if cfg.mu_rule == "radau" else gauss_mu)(int(nmu[m]))
        mus.append(x)
        wfl.append(2.0 * w * x)
    t = lambda v, dt: torch.as_tensor(np.ascontiguousarray(v), dtype=dt, device=dev)
    B = m3drt.B1D_DTYPES
    arrays = dict(nd=t(nd, B["nd"]), nf=t(nf, B["nf"]), nmu=t(nmu, B["nmu"]), nline=t(nline, B["nline"]),
                  d_off=t(d_off, B["d_off"]), l_off=t(l_off, B["l_off"]), ld_off=t(ld_off, B["ld_off"]),
                  mu_off=t(mu_off, B["mu_off"]), f_off=t(f_off, B["f_off"]), i_off=t(i_off, B["i_off"]),
                  lam0=t(lam0, B["lam0"]), dlam=t(dlam, B["dlam"]),
                  dtau=dtau.float().contiguous(), hckT=hckT.float().contiguous(), ivth=(1.0 / vth).float().contiguous(),
                  cont_a=cont_a.float().contiguous(), cont_b=cont_b.float().contiguous(),
                  line_off=t(line_off, B["line_off"]), line_cl=t(line_cl, B["line_cl"]), line_rec=rec.reshape(-1),
                  mu=t(np.concatenate(mus), B["mu"]), wflux=t(np.concatenate(wfl), B["wflux"]),
                  I_top=torch.zeros(int((nmu.astype(np.int64) * nf).sum()), dtype=torch.float32, device=dev),
                  flux=torch.zeros(int(nf.sum()), dtype=torch.float32, device=dev))
    meta = dict(nd=nd, nf=nf, nmu=nmu, nline=nline, d_off=d_off, l_off=l_off, ld_off=ld_off, mu_off=mu_off,
                f_off=f_off, i_off=i_off, lam0=lam0, dlam=dlam, teff=teff, logg=logg)
    return arrays, meta


def updates(meta):
    """(depth, frequency, mu) updates of one synthesis of the batch."""
    return int((meta["nd"].astype(np.int64) * meta["nf"] * meta["nmu"]).sum())


def profile_evaluations(arrays, meta):
    """(line, depth, frequency) profile evaluations of one synthesis: per (line, depth)
    window the frequencies inside it."""
    rec = arrays["line_rec"].view(-1, 4)
    reach = rec[:, 2].double()
    n = len(meta["nd"])
    cnt = torch.as_tensor(meta["nline"].astype(np.int64) * meta["nd"], device=rec.device)
    atm = torch.repeat_interleave(torch.arange(n, device=rec.device), cnt)
    dl = torch.as_tensor(meta["dlam"], device=rec.device)[atm]
    nfa = torch.as_tensor(meta["nf"].astype(np.float64), device=rec.device)[atm]
    inside = torch.where(reach >= 0, torch.minimum(2.0 * reach / dl + 1.0, nfa), torch.zeros_like(reach))
    return float(inside.sum())


def main():
    """Generate the workload, plan it (autotuned), time the persistent graph against
    the plain one-launch-per-atmosphere path, and print one JSON line."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--natm", type=int, default=8192)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--config0", action="store_true", help="the single 72-point, 4000-wavelength atmosphere")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    ctx = m3drt.Context(0)
    t0 = time.perf_counter()
    cfg = CONFIG0 if args.config0 else B1dConfig(natm=args.natm)
    args.natm = cfg.natm
    arrays, meta = make_batch(cfg)
    torch.cuda.synchronize()
    log(f"[b1d] batch of {args.natm} atmospheres in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    plan = ctx.b1d_plan(arrays, LAM_REF)
    torch.cuda.synchronize()
    tile, times = plan.info()
    log(f"[b1d] plan (autotune {times}) -> tile {tile} in {time.perf_counter() - t0:.1f} s")
    plan.run()
    torch.cuda.synchronize()
    ts = []
    for _ in range(args.reps):
        a = time.perf_counter()
        plan.run()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - a)
    F_p = arrays["flux"].clone()
    I_p = arrays["I_top"].clone()
    log(f"[b1d] persistent {min(ts) * 1e3:.1f} ms")
    arrays["flux"].zero_()
    ctx.b1d_synth_plain(arrays, LAM_REF, tile=tile)
    torch.cuda.synchronize()
    a = time.perf_counter()
    ctx.b1d_synth_plain(arrays, LAM_REF, tile=tile)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - a
    log(f"[b1d] plain {t_plain * 1e3:.1f} ms")
    rel = float(((arrays["flux"] - F_p).abs() / F_p.abs().clamp(min=1e-30)).max())
    nu = updates(meta)
    nev = profile_evaluations(arrays, meta)
    t_p = min(ts)
    out = dict(workload="batched_uneven_1d_lte", atmospheres=args.natm, updates=nu, profile_evaluations=nev,
               tile=dict(fpl=tile[0], threads=tile[1]), autotune_ms=times, persistent_ms=t_p * 1e3,
               plain_ms=t_plain * 1e3, speedup=t_plain / t_p, updates_per_s=nu / t_p,
               profile_evaluations_per_s=nev / t_p, flux_max_rel_diff_plain=rel,
               intensity_bitwise_equal=bool(torch.equal(arrays["I_top"], I_p)))
    plan.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
