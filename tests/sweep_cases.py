"""Seeded sweep inputs shared by the oracle and GPU tests.

A case holds the reference-layout work arrays of one directional sweep (planes
0 and nz+1 are ghost planes, as in rt_patch_t%work(f,x,y,0:nz+1,k)) plus the
per-cell chi500 arrays, and converts them into the layouts the C ABI takes.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

DS = (50e5, 50e5, 10e5)  # cm: 50 km horizontally, 10 km vertically


def gauss_rays():
    vec, w = O.gauss_quadrature(2, 6)
    return vec, w


def make_case(nf=8, nx=12, ny=10, nz=9, seed=0, vec=(0.3, -0.4, 0.866), ds=DS, eta=False,
              maser=False, top_ghost_zero=True):
    """Random positive chi/S with a vertical stratification, chi500 and chi500_up
    consistent with a tilted ray, and the upstream ghost plane."""
    rng = np.random.default_rng(seed)
    vec = np.asarray(vec, np.float64)
    vec = vec / np.linalg.norm(vec)
    dir = 1 if vec[2] > 0 else -1
    z = np.arange(nz + 2, dtype=np.float64)[:, None, None]
    # z grows upward; chi falls by ~e^-0.35 per plane, lines multiply it per frequency
    line = 10 ** rng.uniform(0.0, 2.0, nf)
    c500 = 2e-7 * np.exp(-0.35 * (z - 1)) * np.exp(0.15 * rng.standard_normal((nz + 2, ny, nx)))
    chi = c500[..., None] * line * np.exp(0.1 * rng.standard_normal((nz + 2, ny, nx, nf)))
    S = 0.05 * np.exp(-0.04 * z[..., None]) * (1 + 0.2 * rng.random((nz + 2, ny, nx, nf)))
    if maser:
        neg = rng.random(chi.shape) < 0.03
        chi = np.where(neg, -0.3 * chi, chi)
    chi500 = c500[1:-1].astype(np.float32)
    # chi500 at the upwind point: the previous plane (in ray order), perturbed
    if dir > 0:
        prev = c500[0:-2]
    else:
        prev = c500[2:]
    chi500_up = (prev * np.exp(0.02 * rng.standard_normal(prev.shape))).astype(np.float32)
    gz = 0 if dir > 0 else nz + 1
    g500 = c500[gz].astype(np.float32)
    I = np.zeros_like(chi)
    if dir < 0 and top_ghost_zero:
        chi[gz] = 0.0
        S[gz] = 0.0
        I[gz] = 0.0
    else:
        I[gz] = S[gz] * 1.02
    case = dict(chi=chi.astype(np.float32), S=S.astype(np.float32), I=I.astype(np.float32),
                chi500=chi500, chi500_up=chi500_up, g500=g500, dir=dir, vec=vec, ds=ds, gz=gz,
                nf=nf, nx=nx, ny=ny, nz=nz)
    if eta:
        nrow = max(1, (nx * ny * nz) // 3)
        cells = rng.choice(nx * ny * nz, size=nrow, replace=False)
        eta_row = np.zeros(nx * ny * nz, np.int32)
        eta_row[cells] = np.arange(1, nrow + 1, dtype=np.int32)
        case["eta_row"] = eta_row.reshape(nz, ny, nx)
        case["eta_mask"] = (rng.random((nrow, nf)) < 0.4).astype(np.uint8)
    return case


def run_oracle(case, fp64=False, split_rates=False):
    return O.short_char2(case["chi"], case["S"], case["I"], case["chi500"], case["chi500_up"],
                         case["g500"], case["dir"], case["ds"], case["vec"],
                         eta_row=case.get("eta_row"), eta_mask=case.get("eta_mask"),
                         split_rates=split_rates, fp64=fp64)


def interior(a):
    """Planes 1..nz of a work array."""
    return a[1:-1]


def to_native(cube):
    """(nz, ny, nx, nf) reference order -> (nz, nf, ny, nx) native order."""
    return np.ascontiguousarray(np.moveaxis(cube, 3, 1))


def plane_to_native(p):
    """(ny, nx, nf) -> (nf, ny, nx)."""
    return np.ascontiguousarray(np.moveaxis(p, 2, 0))


def max_rel(a, b, floor=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(b), floor)
    m = den > 0
    if not np.any(m):
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)[m] / den[m]))
