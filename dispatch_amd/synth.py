"""Seeded synthetic atmospheres and opacity cubes for the sweep benchmarks.

The 3D LTE formal-solve workload: a 120 x 120 x 230 periodic box (50 km cells
horizontally, 10 km vertically), grey T(tau) with Teff = 5800 K plus a Gaussian
random field (correlation length 8 cells, 8 % rms, tapered to zero below tau = 10),
1024 frequencies equidistant in ln(lambda) over 5000-5004 A with three Voigt lines
(line-centre tau = 1 near log tau500 = -1, -2.5, -4) on a grey + lambda^3
continuum, chi and S = B_nu(T) (occupation units) precomputed per (f, cell) in FP32,
24 rays (2 Gauss mu x 6 azimuths, mirrored; set_gauss_quadratures(2,6) of
SRC/ray_mod.f90).  Seed 20261202 + workload index.

Everything large is built on the GPU with torch, a plane slab at a time; cubes are
in the native layout (nz, nf, ny, nx).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

H_PLANCK = 6.62607015e-27
K_BOLTZ = 1.380649e-16
C_LIGHT = 2.99792458e10
AMU = 1.66053907e-24

# Gauss nodes and weights of set_gauss_quadratures(2, .): default-real literals in
# the reference (SRC/ray_mod.f90:284-286), hence the float32 rounding
_GAUSS2_MU = (float(np.float32(0.78867513)), float(np.float32(0.21132487)))
_GAUSS2_W = (0.5, 0.5)
_PI_REF = 3.14159265359  # HOP/const_mod.f90:14


def gauss_rays(nphi: int = 6):
    """Mirrored Gauss-Legendre set with 2 inclinations (nphi_copies with
    Jnu_weights, SRC/ray_mod.f90:392-467): directions (n, 3) and J weights (n)."""
    vec, wts = [], []
    for mu, wmu in zip(_GAUSS2_MU, _GAUSS2_W):
        th = math.acos(mu)
        for p in range(nphi):
            phi = 2.0 * _PI_REF / nphi * p
            v = [math.sin(th) * math.cos(phi), math.sin(th) * math.sin(phi), mu]
            vec.append([0.0 if abs(c) < 1e-3 else c for c in v])
            wts.append(wmu / nphi)
    n = len(vec)
    vec = vec + [[-c for c in v] for v in vec]
    wts = [w / 2 for w in wts + wts]
    return np.array(vec), np.array(wts)


@dataclass
class Box:
    nx: int
    ny: int
    nz: int
    nf: int
    ds: tuple
    lam: np.ndarray           # (nf,) wavelengths, cm
    T: torch.Tensor           # (nz, ny, nx) K, float64
    chi500: torch.Tensor      # (nz, ny, nx) float32, cm^-1
    vz: torch.Tensor          # (nz, ny, nx) cm/s
    T_ghost_bot: torch.Tensor     # (ny, nx)
    g500_bot: torch.Tensor        # (ny, nx) float32: chi500 one plane below
    g500_top: torch.Tensor        # (ny, nx) float32: chi500 one plane above


def _random_field(shape, corr_cells, rng):
    """Unit-rms Gaussian random field, periodic, FFT filter exp(-k^2 l^2 / 2)."""
    noise = rng.standard_normal(shape)
    ks = [np.fft.fftfreq(n) * 2 * np.pi for n in shape]
    kz, ky, kx = np.meshgrid(*ks, indexing="ij")
    filt = np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) * corr_cells ** 2 / 2)
    f = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    return f / f.std()


def make_box(nx=120, ny=120, nz=230, nf=1024, seed=20261203, device="cuda",
             dz_km=10.0, dx_km=50.0, lam0_A=5000.0, band_A=4.0):
    """The atmosphere of the 3D LTE formal-solve workload."""
    rng = np.random.default_rng(seed)
    dz = dz_km * 1e5
    ds = (dx_km * 1e5, dx_km * 1e5, dz)
    # grey stratification: log tau500 from +2 (bottom plane) to -5 (top plane)
    ltau = np.linspace(2.0, -5.0, nz)
    tau = 10 ** ltau
    teff = 5800.0
    T1 = teff * (0.75 * (tau + 2.0 / 3.0)) ** 0.25
    H = (nz - 1) * dz / ((ltau[0] - ltau[-1]) * math.log(10.0))   # tau = chi500 * H
    # temperature fluctuations, tapered to zero below tau = 10
    dT = _random_field((nz, ny, nx), 8.0, rng)
    taper = 0.5 * (1.0 - np.tanh((ltau - 1.0) / 0.2))
    T = T1[:, None, None] * (1.0 + 0.08 * dT * taper[:, None, None])
    # chi500 follows the hydrostatic exponential and the H- sensitivity to T
    c500 = (tau / H)[:, None, None] * np.exp(2.0 * (T / T1[:, None, None] - 1.0))
    vz = 2e5 * _random_field((nz, ny, nx), 4.0, rng)
    lam = lam0_A * 1e-8 * np.exp(np.linspace(0.0, math.log((lam0_A + band_A) / lam0_A), nf))
    Tg = 2 * T[0] - T[1]
    g_bot = c500[0] ** 2 / c500[1]
    g_top = c500[-1] ** 2 / c500[-2]
    tt = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
    return Box(nx, ny, nz, nf, ds, lam, tt(T), tt(c500, torch.float32), tt(vz), tt(Tg),
               tt(g_bot, torch.float32), tt(g_top, torch.float32))


def _voigt(a, v):
    """Voigt function H(a, v) by the pseudo-Voigt of Thompson, Cox and Hastings
    (1987): adequate for synthetic line shapes (not a reference quantity)."""
    fg = 2.0 * math.sqrt(math.log(2.0))                 # Gaussian FWHM in Doppler units
    fl = 2.0 * a
    f = (fg ** 5 + 2.69269 * fg ** 4 * fl + 2.42843 * fg ** 3 * fl ** 2 + 4.47163 * fg ** 2 * fl ** 3
         + 0.07842 * fg * fl ** 4 + fl ** 5) ** 0.2
    r = fl / f
    eta = 1.36603 * r - 0.47719 * r ** 2 + 0.11116 * r ** 3
    g = torch.exp(-4 * math.log(2.0) * (v / f) ** 2) * (2 * math.sqrt(math.log(2.0) / math.pi) / f)
    l = (2 / math.pi) * (f / 2) / (4 * v ** 2 + f ** 2) * 2
    phi = eta * l + (1 - eta) * g                       # normalised in v
    return phi * math.sqrt(math.pi)                     # H(a, v) normalisation: H(0, 0) = 1


LINES = (  # (centre offset in A from lam0, line-to-continuum ratio at the centre, damping a, mass amu)
    (1.0, 10 ** 1.0, 0.02, 56.0),
    (2.0, 10 ** 2.5, 0.03, 40.0),
    (3.0, 10 ** 4.0, 0.05, 24.0),
)


def planck_occupation(lam, T):
    """B_nu in occupation units, 1 / (exp(h nu / kT) - 1), as
    lte_background_opacities forms it (SRC/rt_patch_mod.f90:2004-2023)."""
    x = (H_PLANCK * C_LIGHT / (lam * K_BOLTZ)) / T
    beta = torch.clamp(torch.exp(-x), min=np.finfo(np.float32).tiny)
    return beta / (1.0 - beta)


def opacity_source(box: Box, z0: int, z1: int, lam0_A=5000.0):
    """chi and S (float64) for planes z0..z1-1, native layout (nz, nf, ny, nx)."""
    dev = box.T.device
    lam = torch.as_tensor(box.lam, dtype=torch.float64, device=dev)[None, :, None, None]
    T = box.T[z0:z1][:, None]
    c5 = box.chi500[z0:z1].double()[:, None]
    vz = box.vz[z0:z1][:, None]
    cont = 0.7 + 0.3 * (lam / (lam0_A * 1e-8)) ** 3
    eta = torch.zeros((z1 - z0, box.nf, box.ny, box.nx), dtype=torch.float64, device=dev)
    for off, ratio, a, mass in LINES:
        lc = (lam0_A + off) * 1e-8 * (1.0 + vz / C_LIGHT)
        vth = torch.sqrt(2 * K_BOLTZ * T / (mass * AMU) + 1e5 ** 2)
        dl = lc * vth / C_LIGHT
        v = (lam - lc) / dl
        strength = ratio * (T / 5800.0) ** -3
        eta = eta + strength * _voigt(a, v)
    chi = c5 * (cont + eta)
    S = planck_occupation(lam, T)
    return chi, S


def chi500_upwind(box: Box, vec):
    """chi500 at the upwind point of every cell for direction vec: the model's
    chi500 on the previous plane in ray order (the ghost plane's for the first),
    interpolated bilinearly in its log at the displaced position."""
    dsx, dsy, dsz = box.ds
    dsup = dsz / abs(vec[2])
    wx = -vec[0] / dsx * dsup
    wy = -vec[1] / dsy * dsup
    lc = torch.log(box.chi500.double())
    if vec[2] > 0:
        prev = torch.cat([torch.log(box.g500_bot.double())[None], lc[:-1]], 0)
    else:
        prev = torch.cat([lc[1:], torch.log(box.g500_top.double())[None]], 0)

    def shift(a, d, dim):
        i0 = math.floor(d)
        fr = d - i0
        a0 = torch.roll(a, -i0, dims=dim)
        a1 = torch.roll(a, -(i0 + 1), dims=dim)
        return (1 - fr) * a0 + fr * a1

    out = shift(shift(prev, wx, 2), wy, 1)
    return torch.exp(out).float().contiguous()


def build_cubes(box: Box, dtype=torch.float32, slab=23):
    """chi and S cubes (nz, nf, ny, nx) in dtype, built slab by slab; FP64 cubes
    are the FP32-rounded values widened, so both sweeps see the same inputs."""
    shape = (box.nz, box.nf, box.ny, box.nx)
    chi = torch.empty(shape, dtype=dtype, device=box.T.device)
    S = torch.empty(shape, dtype=dtype, device=box.T.device)
    for z0 in range(0, box.nz, slab):
        z1 = min(box.nz, z0 + slab)
        c, s = opacity_source(box, z0, z1)
        chi[z0:z1] = c.float().to(dtype)
        S[z0:z1] = s.float().to(dtype)
        del c, s
    return chi, S


def ghost_planes(box: Box, chi_cube, dtype):
    """Upstream ghost planes (chi, S, I), native (nf, ny, nx): the LTE plane below
    the bottom (set_upstream_boundary / lte_ghost, SRC/rt_patch_mod.f90:3352-3373:
    chi = chi(z=1)*g500/chi500(1), S = I = B_nu(T_ghost)) for upward rays, and
    zeros above the top for downward rays."""
    dev = box.T.device
    lam = torch.as_tensor(box.lam, dtype=torch.float64, device=dev)[:, None, None]
    r500 = (box.g500_bot.double() / box.chi500[0].double())[None]
    chi_b = (chi_cube[0].double() * r500).float().to(dtype)
    S_b = planck_occupation(lam, box.T_ghost_bot[None]).float().to(dtype)
    zero = torch.zeros_like(chi_b)
    return dict(up=(chi_b.contiguous(), S_b.contiguous(), S_b.clone()),
                down=(zero, zero.clone(), zero.clone()))


# ------------------------------------------------------------------ model atom

EV_K = 11604.518  # K per eV
SAHA = 2.4146830e15  # (2 pi m_e k / h^2)^(3/2) [cm^-3 K^-3/2]


@dataclass
class Atom:
    """A synthetic model atom: nk-1 bound levels of one ion plus the ground level
    of the next ion (the continuum).  Transitions are lines first, then continua;
    levels and transitions are 1-based as in the reference's atom files."""
    nk: int
    ev: np.ndarray        # (nk) level energies [eV]; the last is the ionisation limit
    g: np.ndarray         # (nk) statistical weights
    ion: np.ndarray       # (nk) ion stage (1 bound, 2 continuum)
    ti: np.ndarray        # (nrad) lower level
    tj: np.ndarray        # (nrad) upper level
    nline: int
    f: np.ndarray         # (nline) oscillator strengths
    Aji: np.ndarray       # (nline) [1/s]
    nu0: np.ndarray       # (nrad) transition frequency [Hz]
    sigma0: np.ndarray    # (ncont) photo-ionisation cross section at the edge [cm^2]
    pi: np.ndarray        # (npair) collisional pairs
    pj: np.ndarray

    @property
    def nrad(self):
        return len(self.ti)


def make_atom(nbound=20, nline=80, ncont=20, seed=20261204, chi_ion=7.6):
    """20 bound levels with energies sorted uniform in 0-7 eV plus the continuum; 80
    bound-bound pairs with f log-uniform in 1e-4..1, 20 bound-free with hydrogenic
    nu^-3 cross sections of edge value 1e-18 cm^2."""
    rng = np.random.default_rng(seed)
    ev = np.concatenate([[0.0], np.sort(rng.uniform(0.0, 7.0, nbound - 1)), [chi_ion]])
    g = np.concatenate([rng.integers(1, 10, nbound).astype(np.float64), [1.0]])
    ion = np.concatenate([np.ones(nbound, np.int32), [2]])
    pairs = [(i, j) for i in range(1, nbound + 1) for j in range(i + 1, nbound + 1)]
    sel = rng.choice(len(pairs), size=nline, replace=False)
    lines = sorted(pairs[s] for s in sel)
    ti = [p[0] for p in lines] + list(range(1, ncont + 1))
    tj = [p[1] for p in lines] + [nbound + 1] * ncont
    f = 10 ** rng.uniform(-4, 0, nline)
    nu0 = np.array([(ev[j - 1] - ev[i - 1]) * 2.417989242e14 for i, j in zip(ti, tj)])
    gi = g[np.array(ti[:nline]) - 1]
    gj = g[np.array(tj[:nline]) - 1]
    Aji = 6.670e15 * gi / gj * f / (C_LIGHT / nu0[:nline] * 1e8) ** 2
    pi_, pj_ = [], []
    for i, j in zip(ti, tj):
        pi_.append(i)
        pj_.append(j)
    return Atom(nbound + 1, ev, g, ion, np.array(ti, np.int32), np.array(tj, np.int32), nline, f, Aji,
                nu0, np.full(ncont, 1e-18), np.array(pi_, np.int32), np.array(pj_, np.int32))


def lte_populations(atom: Atom, T, ne, totn):
    """Saha-Boltzmann LTE populations (nk, ncell) normalised to totn."""
    T = np.asarray(T, np.float64)
    kT = T / EV_K
    bound = atom.ion == 1
    w = np.empty((atom.nk, T.size))
    w[bound] = atom.g[bound, None] * np.exp(-atom.ev[bound, None] / kT[None])
    # next ion's ground, relative to the neutral ground (g = g_ground of the ion)
    chi = atom.ev[-1]
    w[-1] = (atom.g[-1] * 2.0 / atom.g[0]) * SAHA * T ** 1.5 * np.exp(-chi / kT) / ne * atom.g[0]
    w *= totn / w.sum(axis=0)
    return w


def van_regemorter(atom: Atom, T, ne, nstar, ion_scale=1e-2):
    """Upward collision rates (npair, ncell): van Regemorter for the lines
    (14.5 Allen Ryd/kT f max(0.2 e^-u, sqrt3/2pi E1(u)) / dE, SRC/atom_mod.f90:
    1790-1838 with the recipe's constants) and a Seaton-type ionisation rate
    1.55e13 ne T^-1/2 s e^-u / u.  Seaton's s is gbar sigma0 (0.1 times the edge
    cross section in cm^2, ~1e-19); the default s = 1e-2 makes the continuum
    strongly collisional, which keeps the statistical-equilibrium test cases, with
    their randomised radiative rates, away from negative populations."""
    from scipy.special import exp1
    T = np.asarray(T, np.float64)
    sqt = np.sqrt(T)
    e_kT = EV_K / T
    rydk, e_k = 157887.51, 11604.518
    c0 = np.pi * (0.529177e-8) ** 2 * np.sqrt(8 * K_BOLTZ / 9.1093837e-28 / np.pi)
    allen = c0 * rydk * ne / sqt
    regem = 14.5 * allen * rydk / e_k
    seaton = 1.55e13 * ne / sqt
    out = np.zeros((len(atom.pi), T.size))
    for q, (i, j) in enumerate(zip(atom.pi, atom.pj)):
        dev = max(atom.ev[j - 1] - atom.ev[i - 1], 1e-8)
        u0 = dev * e_kT
        ex = np.exp(-u0)
        if q < atom.nline:
            f = max(atom.f[q], 1e-6 * atom.g[j - 1])
            out[q] = regem * f * np.maximum(0.2 * ex, np.sqrt(3) / (2 * np.pi) * exp1(u0)) / dev
        else:
            out[q] = seaton * ion_scale / u0 * ex
    return out
