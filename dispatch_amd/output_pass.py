"""The output pass of the 3D LTE spectrum: radiative_transfer's final branch
(SRC/rt_patch_mod.f90:1669-1691) over the output rays.

For each upward output ray the work arrays are rotated so that the rays become
columns (rotate_array, SRC/rt_patch_mod.f90:3809-3863, pivoting about the top
plane), a column solver runs from the ghost plane below the bottom, built from
the unrotated model (the LTE ghost for the hermite solvers, the PCHIP one for the
others) (long_char_linear, long_char_quadratic, long_char_hermite or
long_char_hermite2, SRC/rt_solvers_mod.f90:666-757, 891-1020, 1052-1504,
1536-1895), and the exit planes give
save_intensity's emergent products ie and flux (SRC/rt_patch_mod.f90:4151-4195,
dispatch_amd/save.py).  long_char_hermite needs the thermodynamic state along the
ray (the reference admits it only with ray-aligned compositions): T and Pe are
rotated with chi and S, beta is formed from the rotated T, and dPe/dz, dlnT/dz are
central differences of the rotated fields.

Workload (python -m dispatch_amd.output_pass): the 120 x 120 x 230 box of the 3D
LTE formal-solve workload (synthetic, seeded), its 12 upward rays (2 Gauss mu x 6
azimuths), frequency windows of nf points.  Pe is synthetic: proportional to the
continuum opacity chi500 (H- opacity scales with Pe), 10 dyn/cm^2 at tau500 = 1.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np
import torch

from . import m3drt, synth
from .constants import CC, H_K
from .save import emergent_products

SOLVERS = ("long_linear", "quadratic", "hermite", "hermite2")


def rotate(ctx, x, vec, ds, bot2pivot):
    out = torch.empty_like(x)
    ctx.rotate_array(x, out, vec, ds, bot2pivot, layout="native")
    return out


def dz_gradient(a, dz):
    """d/dz of (nz, ny, nx) by central differences, one-sided at the ends [per cm]."""
    g = torch.empty_like(a)
    g[1:-1] = (a[2:] - a[:-2]) / (2 * dz)
    g[0] = (a[1] - a[0]) / dz
    g[-1] = (a[-1] - a[-2]) / dz
    return g


def ghost_planes(ctx, box, chi, S, pe):
    """The bottom ghost planes of the output pass, built before the cubes are rotated
    (set_upstream_boundary precedes radiative_transfer's rotate_array,
    SRC/rt_patch_mod.f90:1034-1036), from the unrotated model: the LTE ghost for the
    hermite solvers (lte_ghost_solver, SRC/rt_params_mod.f90:78-85; chi(first plane)
    g500/chi500(first plane), S = I = B_nu(T below the bottom)), the PCHIP ghost for the
    others (chi and S extrapolated from the first planes, clamped at 0, I = S), and the
    model below the bottom the hermite solvers take (bottom_ghost_t: T, Pe, dPe/dz,
    dlnT/dz)."""
    nf, ny, nx = chi.shape[1:]
    mk = lambda: [torch.empty((nf, ny, nx), dtype=torch.float32, device=chi.device) for _ in range(3)]
    nu = torch.as_tensor(CC / box.lam, dtype=torch.float64, device=chi.device)
    lte, pchip = mk(), mk()
    ctx.set_upstream_boundary("lte", lte, chi1=chi[0].contiguous(), g500=box.g500_bot, chi500_1=box.chi500[0],
                              T_ghost=box.T_ghost_bot.float().contiguous(), nu=nu)
    ctx.set_upstream_boundary("pchip", pchip, chi1=chi[0].contiguous(), S1=S[0].contiguous(),
                              chi2=chi[1].contiguous(), S2=S[1].contiguous(), ds3=box.ds[2])
    dz = box.ds[2]
    dpe0 = (pe[1] - pe[0]) / dz
    Tb = box.T.float()
    dlnt0 = (torch.log(Tb[1]) - torch.log(Tb[0])) / dz
    model = dict(pe=torch.clamp(pe[0] - dz * dpe0, min=1e-3).contiguous(), dpe_dz=dpe0.contiguous(),
                 temp=box.T_ghost_bot.float().contiguous(), dlnt_dz=dlnt0.contiguous())
    return dict(lte=lte, pchip=pchip, model=model)


def ghost_of(solver, ghosts):
    return ghosts["lte"] if solver in ("hermite", "hermite2") else ghosts["pchip"]


def ray_state(ctx, box, chi, S, pe, vec, ghosts):
    """The rotated cubes and fields of one ray; native layout (nz, nf, ny, nx)."""
    nz = box.nz
    dz = box.ds[2]
    top = nz * dz
    chi_r = rotate(ctx, chi[None], vec, box.ds, top)[0]
    S_r = rotate(ctx, S[None], vec, box.ds, top)[0]
    cell = torch.stack([box.T.float(), pe])[:, :, None]                  # (2, nz, 1, ny, nx)
    cell = rotate(ctx, cell.contiguous(), vec, box.ds, top)[:, :, 0]
    T, pe_r = cell[0].contiguous(), cell[1].contiguous()
    nu = torch.as_tensor(CC / box.lam, dtype=torch.float64, device=chi.device)
    hknu = (H_K * nu).float()                       # real(h_k nu, 4) (SRC/rt_solvers_mod.f90:1096)
    # beta = max(exp(-h nu / k T), tiny) of lte_background_opacities on the rotated
    # composition, plane by plane into the native layout
    nf = nu.numel()
    beta = torch.empty_like(chi_r)
    B = torch.empty((nf, box.ny * box.nx), dtype=torch.float32, device=chi.device)
    T64 = T.double()
    for z in range(nz):
        ctx.lte_background(nu, T64[z].reshape(-1), beta[z], B)
    dpe = dz_gradient(pe_r, dz)
    dlnt = dz_gradient(torch.log(T), dz)
    fields = dict(beta=beta, pe4=pe_r, temp=T, dpe_dz=dpe.contiguous(), dlnt_dz=dlnt.contiguous(),
                  hknu=hknu.contiguous())
    return chi_r, S_r, fields


def solve(ctx, solver, chi, S, fields, ghosts, box, vec, exit_I):
    kw = dict(dir=1, ds=box.ds, vec=vec, layout="native", exit_I=exit_I)
    gh = ghost_of(solver, ghosts)
    if solver in ("hermite", "hermite2"):
        ctx.lc_sweep_hermite(chi, S, *gh, fields=fields, ghost="bottom", ghost_fields=ghosts["model"],
                             hybrid=solver == "hermite2", **kw)
    else:
        ctx.sc_sweep(chi, S, *gh, None, None, None, solver=solver, **kw)


def synthetic_pe(box):
    c5 = box.chi500.float()
    ref = torch.quantile(c5.flatten()[:: max(1, c5.numel() // 100000)], 0.5)
    return (10.0 * c5 / ref).contiguous()


def run(nf=256, windows=1, solvers=SOLVERS, rays=None, device=0, save_dir=None, reps=1):
    """Emergent spectra of every solver over the upward rays; per-kernel timings."""
    torch.cuda.set_device(device)
    ctx = m3drt.Context(device)
    vecs, wts = synth.gauss_rays(6)
    up = [i for i in range(len(vecs)) if vecs[i][2] > 0]
    if rays is not None:
        up = up[:rays]
    # flux weights of the output rays: 2 pi sum wmu mu I over the upward hemisphere
    wflux = [float(wts[i] * 2 * vecs[i][2]) for i in up]
    tim = {k: [] for k in ("rotate",) + tuple(solvers)}
    ie_all, flux_all = {s: [] for s in solvers}, {s: [] for s in solvers}
    nbytes_cube = None
    for w in range(windows):
        box = synth.make_box(nf=nf, seed=20261203, lam0_A=5000.0 + 4.0 * w, device=f"cuda:{device}")
        chi, S = synth.build_cubes(box)
        pe = synthetic_pe(box)
        ghosts = ghost_planes(ctx, box, chi, S, pe)
        nbytes_cube = chi.numel() * 4
        exits = {s: torch.zeros((len(up), nf, box.ny, box.nx), dtype=torch.float32, device=chi.device)
                 for s in solvers}
        for j, i in enumerate(up):
            vec = [float(v) for v in vecs[i]]
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            state = ray_state(ctx, box, chi, S, pe, vec, ghosts)
            e1.record()
            torch.cuda.synchronize()
            tim["rotate"].append(e0.elapsed_time(e1))
            for s in solvers:
                for _ in range(reps):
                    e0.record()
                    solve(ctx, s, *state, ghosts, box, vec, exits[s][j])
                    e1.record()
                    torch.cuda.synchronize()
                    tim[s].append(e0.elapsed_time(e1))
            del state
        nu = CC / box.lam
        for s in solvers:
            ie, flux, _ = emergent_products(exits[s], nu, wflux)
            ie_all[s].append(ie)
            flux_all[s].append(flux)
            if save_dir is not None:
                from .save import save_intensity
                save_intensity(f"{save_dir}/{s}", exits[s], nu, wflux, key=(1, w + 1))
        del chi, S, exits
    ctx.close()
    # per (f, cell): chi and S read (8 B; the Hermite sweep also beta, 12 B); the exit
    # plane written once per column
    rd = dict(long_linear=8, quadratic=8, hermite=12, hermite2=8)
    out = dict(workload=f"output pass: 120x120x230 box, {len(up)} upward rays, {windows} window(s) of {nf} f",
               ms_per_ray={k: float(np.median(v)) for k, v in tim.items()})
    out["GB/s"] = {s: rd[s] * nbytes_cube / 4 / (out["ms_per_ray"][s] * 1e-3) / 1e9 for s in solvers}
    flux = {s: np.concatenate([f.reshape(-1) for f in flux_all[s]]) for s in solvers}
    out["flux_mean"] = {s: float(flux[s].mean()) for s in solvers}
    if "hermite" in solvers and "quadratic" in solvers:
        out["flux_rel_diff_hermite_quadratic"] = float(np.max(np.abs(flux["hermite"] / flux["quadratic"] - 1)))
    return out


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--nf", type=int, default=256)
    p.add_argument("--windows", type=int, default=1)
    p.add_argument("--rays", type=int, default=None)
    p.add_argument("--reps", type=int, default=1)
    p.add_argument("--solver", choices=SOLVERS + ("all",), default="all")
    p.add_argument("--save", default=None, help="write ie/flux per solver under this run directory")
    a = p.parse_args(argv)
    solvers = SOLVERS if a.solver == "all" else (a.solver,)
    t0 = time.time()
    out = run(a.nf, a.windows, solvers, a.rays, save_dir=a.save, reps=a.reps)
    out["wall_s"] = time.time() - t0
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
