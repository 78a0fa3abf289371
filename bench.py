#!/usr/bin/env python3
"""Benchmark of the 3D short-characteristics formal solution on one MI355X.

Workload: the 3D LTE formal solve on a 120 x 120 x 230 periodic box (DISPATCH
slabs of 23 planes, swept as one column), 24 rays (2 Gauss mu x 6 azimuths,
mirrored) x 1024 frequencies, chi and S precomputed per (f, cell); see
dispatch_amd/synth.py.  One step = one formal solution: the 24 directional sweeps
(short_char2 semantics) with J accumulated over the rays and the emergent
intensity of the 12 upward rays written out.  Inputs are resident in HBM before
the timed region.

Prints ONE JSON line: FP32 updates/s (the metric), the plain FP64 sweep beside it,
the FP32 error against FP64, the roofline of the FP32 sweep kernel (HIP events
around every sweep launch on its stream) and the CPU baseline (the oracle's
restatement with OpenMP over frequency windows, on a bounded sample).

    python bench.py [--steps K] [--warmup W] [--no-fp64] [--no-cpu] [--calib-copy]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ray·freq·cell updates/s FP32 vs plain FP64 at ≤1e-5 rel err; adjoint/fwd time"
UNIT = "ray·freq·cell updates/s"
HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {"f32": 16, "f64": 32}   # chi, S read; J read and written


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpus_visible() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_cores() -> int:
    """Threads for the CPU baseline: OMP_NUM_THREADS where the host sets it (a GPU
    box shares its CPUs between its GPUs and sets it to its share), else every CPU
    this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return min(int(env), cpus_visible())
    return cpus_visible()


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-fp64", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--calib-copy", action="store_true",
                   help="run a plain 4 GiB device copy first (HBM-counter calibration)")
    p.add_argument("--no-adjoint", action="store_true")
    p.add_argument("--no-batch1d", action="store_true")
    p.add_argument("--no-output-pass", action="store_true")
    p.add_argument("--no-adjoint-fd", action="store_true",
                   help="skip the check of the abundance gradients against FP64 central differences")
    p.add_argument("--nf", type=int, default=1024)
    p.add_argument("--nz", type=int, default=230)
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-nlte", action="store_true")
    p.add_argument("--nlte-fp64", action="store_true",
                   help="also run the NLTE iteration with the plain FP64 sweep and compare b")
    p.add_argument("--nlte-max-iter", type=int, default=60)
    p.add_argument("--no-voigt", action="store_true")
    p.add_argument("--no-sweep", action="store_true",
                   help="skip the formal solution (profiling another leg on its own; no metric value)")
    return p.parse_args()


class Solver:
    """The 24-direction formal solution over a resident box in one precision."""

    def __init__(self, ctx, box, dtype, rays, weights, cu):
        import torch
        from dispatch_amd import synth
        self.torch = torch
        self.ctx, self.box, self.dtype = ctx, box, dtype
        self.rays, self.weights, self.cu = rays, weights, cu
        t0 = time.time()
        self.chi, self.S = synth.build_cubes(box, dtype)
        self.ghost = synth.ghost_planes(box, self.chi, dtype)
        self.J = torch.zeros_like(self.chi)
        nup = int((rays[:, 2] > 0).sum())
        self.emergent = torch.zeros((nup, box.nf, box.ny, box.nx), dtype=dtype, device="cuda")
        self.ev = []
        log(f"[bench] {dtype} cubes built in {time.time() - t0:.1f}s")

    def step(self, record=False):
        box = self.box
        self.J.zero_()
        iup = 0
        for r, v in enumerate(self.rays):
            up = v[2] > 0
            gchi, gS, gI = self.ghost["up" if up else "down"]
            g500 = box.g500_bot if up else box.g500_top
            exit_I = None
            if up:
                exit_I = self.emergent[iup]
                iup += 1
            if record:
                e0 = self.torch.cuda.Event(enable_timing=True)
                e1 = self.torch.cuda.Event(enable_timing=True)
                e0.record()
            self.ctx.sc_sweep(self.chi, self.S, gchi, gS, gI, box.chi500, self.cu[r], g500,
                              dir=1 if up else -1, ds=box.ds, vec=v, wj=float(self.weights[r]),
                              J=self.J, exit_I=exit_I)
            if record:
                e1.record()
                self.ev.append((e0, e1))

    def launch_ms(self):
        return [a.elapsed_time(b) for a, b in self.ev]

    def free(self):
        del self.chi, self.S, self.J, self.ghost


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "sc2_fast_f32_box120x120x230_nf1024.md")


def measured_traffic(box):
    """HBM bytes per sweep launch measured by the FETCH_SIZE / WRITE_SIZE counter
    passes (calibrated on a plain copy) recorded in profiles/, for the workload they
    were taken on; None for any other box."""
    if (box.nx, box.ny, box.nz, box.nf) != (120, 120, 230, 1024) or not os.path.exists(TRAFFIC_PROFILE):
        return None, None
    val = kern = None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic_bytes_per_launch:"):
            val = float(line.split(":", 1)[1])
        if line.startswith("kernel:"):
            kern = line.split(":", 1)[1].strip()
    return val, kern


def timed(solver, steps, warmup):
    torch = solver.torch
    for _ in range(warmup):
        solver.step()
    torch.cuda.synchronize()
    solver.ev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.step(record=True)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def max_rel_err(a, b, floor_frac=0.0):
    """max |a - b| / max(|b|, floor_frac*max|b|) in chunks on the device."""
    torch = __import__("torch")
    bmax = float(b.abs().max())
    floor = floor_frac * bmax
    worst = 0.0
    for i in range(a.shape[0]):
        d = (a[i].double() - b[i].double()).abs()
        den = b[i].double().abs().clamp_min(max(floor, 1e-300))
        worst = max(worst, float((d / den).max()))
    return worst


def cpu_baseline(box, chi32, S32, rays, weights, cu, seconds):
    """The oracle's short_char2 + integrate_intensity (OpenMP over frequency
    windows) on the host cores, on a sample of the workload's frequencies."""
    ncores = host_cores()
    import numpy as np
    from oracle import oracle as O
    nfs = min(box.nf, 64)
    f0 = box.nf // 2 - nfs // 2
    nz, ny, nx = box.nz, box.ny, box.nx
    # reference layout (f fastest) with the ghost planes: (nz+2, ny, nx, nfs)
    chi = np.zeros((nz + 2, ny, nx, nfs), np.float32)
    S = np.zeros_like(chi)
    chi[1:-1] = chi32[:, f0:f0 + nfs].permute(0, 2, 3, 1).cpu().numpy()
    S[1:-1] = S32[:, f0:f0 + nfs].permute(0, 2, 3, 1).cpu().numpy()
    c500 = box.chi500.cpu().numpy()
    from dispatch_amd import synth
    gh = synth.ghost_planes(box, chi32, chi32.dtype)
    J = np.zeros((nz, ny, nx, nfs), np.float32)
    done, upd, t_used = 0, 0, 0.0
    for r, v in enumerate(rays):
        up = v[2] > 0
        gchi, gS, gI = gh["up" if up else "down"]
        I = np.zeros_like(chi)
        gz = 0 if up else nz + 1
        chi[gz] = gchi[f0:f0 + nfs].permute(1, 2, 0).cpu().numpy()
        S[gz] = gS[f0:f0 + nfs].permute(1, 2, 0).cpu().numpy()
        I[gz] = gI[f0:f0 + nfs].permute(1, 2, 0).cpu().numpy()
        g500 = (box.g500_bot if up else box.g500_top).cpu().numpy()
        cur = cu[r].cpu().numpy()
        t0 = time.perf_counter()
        O.sweep_windows(chi, S, I, c500, cur, g500, 1 if up else -1, box.ds, v,
                        np.float32(weights[r]), J, nwin=4 * ncores)
        t_used += time.perf_counter() - t0
        done += 1
        upd += nfs * nx * ny * nz
        if t_used >= seconds:
            break
    return dict(value=upd / t_used, unit=UNIT, cores=ncores, kind="port", cpu_model=cpu_model(),
                cpus_visible=cpus_visible(),
                sample=f"{done} of {len(rays)} rays x {nfs} of {box.nf} frequencies on the full "
                       f"{nx}x{ny}x{nz} box ({upd:.3g} updates, {t_used:.1f} s); oracle short_char2 "
                       f"FP32 + integrate_intensity, OpenMP over {4 * ncores} frequency windows")


def adjoint_leg(ctx, args):
    """Gradient / forward cost on the abundance-fit workload (96 x 96 x 160, 24 rays,
    2048 frequencies, 12 lines of 4 elements).  Forwards, one launch per ray and
    nothing kept: the 12 upward rays' exit planes the flux loss needs, and the
    24-ray formal solution with J.  Gradient of the flux chi^2: the checkpointed
    forward of the 12 upward rays (patches of 23 planes, exit planes kept), the
    backward (each patch recomputed from its checkpoint and swept in reverse), and
    the chain to the 4 log abundances.  Checked against FP64 central differences
    (h = 1e-3 dex) unless --no-adjoint-fd."""
    import torch
    from dispatch_amd.fit import FitProblem
    t0 = time.time()
    torch.cuda.reset_peak_memory_stats()
    p = FitProblem(ctx)
    p.make_target()
    log(f"[bench] fit problem ready in {time.time() - t0:.1f}s")
    d0 = [0.0, 0.0, 0.0, 0.0]
    chi, S = p.cubes(d0)
    full, flux = [], []
    with torch.no_grad():
        for i in range(1 + max(1, args.steps // 2)):
            torch.cuda.synchronize()
            a = time.perf_counter()
            p.forward_full(chi, S)
            torch.cuda.synchronize()
            b = time.perf_counter()
            p.forward_flux(chi, S)
            torch.cuda.synchronize()
            if i:
                full.append(b - a)
                flux.append(time.perf_counter() - b)
    del chi, S
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    tms = []
    for i in range(2):
        tm = {}
        loss, g, _ = p.loss_and_grad(d0, timing=tm)
        tms.append(tm)
    tm = min(tms, key=lambda t: t["forward_s"] + t["adjoint_s"] + t["chain_s"])
    t_flux, t_full = min(flux), min(full)
    nup = len(p.up_rays())
    grad_s = tm["forward_s"] + tm["adjoint_s"] + tm["chain_s"]
    out = {"workload": "abundance_fit", "box": [p.box.nx, p.box.ny, p.box.nz], "rays": len(p.vec),
           "frequencies": p.box.nf, "lines": 12, "elements": 4, "patch_planes": p.rs.patch,
           "forward_flux_ms": t_flux * 1e3, "forward_flux_rays": nup,
           "forward_full_ms": t_full * 1e3, "forward_full_rays": len(p.vec),
           "gradient_ms": grad_s * 1e3, "checkpointed_forward_ms": tm["forward_s"] * 1e3,
           "backward_ms": tm["adjoint_s"] * 1e3, "abundance_chain_ms": tm["chain_s"] * 1e3,
           "gradient_over_forward": grad_s / t_flux,
           "backward_over_forward": tm["adjoint_s"] / t_flux,
           "gradient_over_full_forward": grad_s / t_full,
           "per_ray_ms": {"forward": t_flux * 1e3 / nup, "backward": tm["adjoint_s"] * 1e3 / nup},
           "peak_hbm_GB_gradient": torch.cuda.max_memory_allocated() / 1e9,
           "loss": loss, "grad_log_abundance": [float(v) for v in g]}
    log(f"[bench] gradient {grad_s * 1e3:.1f} ms (backward {tm['adjoint_s'] * 1e3:.1f}) vs forward "
        f"{t_flux * 1e3:.1f} ms ({nup} rays)")
    if not args.no_adjoint_fd:
        fd = []
        for e in range(4):
            dp, dm = list(d0), list(d0)
            dp[e] += 1e-3
            dm[e] -= 1e-3
            fd.append((float(p.loss_of_flux(p.flux_fp64(dp))) - float(p.loss_of_flux(p.flux_fp64(dm)))) / 2e-3)
        out["fd_fp64_h1e-3"] = fd
        out["fd_max_rel_err"] = max(abs(float(a) - b) / abs(b) for a, b in zip(g, fd))
        out["fd_tolerance"] = 1e-3
        log(f"[bench] adjoint vs FP64 central differences: {out['fd_max_rel_err']:.2e}")
    del p
    return out


def batch1d_leg(ctx, args):
    """The batched uneven 1D workload (8192 plane-parallel atmospheres, one persistent
    graph-captured launch with the autotuned tile), and against the plain
    one-launch-per-atmosphere path on its first 512 atmospheres (a sample: the plain
    path takes about 20 s for all 8192)."""
    import torch
    from dispatch_amd import batch1d as B
    t0 = time.time()
    arrays, meta = B.make_batch(B.B1dConfig())
    plan = ctx.b1d_plan(arrays, B.LAM_REF)
    tile, times = plan.info()
    log(f"[bench] batch1d ready in {time.time() - t0:.1f}s, tile {tile}")
    plan.run()
    torch.cuda.synchronize()
    ts = []
    for _ in range(max(2, args.steps)):
        a = time.perf_counter()
        plan.run()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - a)
    t_full = min(ts)
    plan.close()
    sub = B.B1dConfig(natm=512)
    sa, sm = B.make_batch(sub)
    sp = ctx.b1d_plan(sa, B.LAM_REF, tile=tile)
    sp.run()
    torch.cuda.synchronize()
    a = time.perf_counter()
    sp.run()
    torch.cuda.synchronize()
    t_sub = time.perf_counter() - a
    F = sa["flux"].clone()
    ctx.b1d_synth_plain(sa, B.LAM_REF, tile=tile)
    torch.cuda.synchronize()
    a = time.perf_counter()
    ctx.b1d_synth_plain(sa, B.LAM_REF, tile=tile)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - a
    same = bool(torch.equal(F, sa["flux"]))
    sp.close()
    nu = B.updates(meta)
    out = {"workload": "batched_uneven_1d_lte", "atmospheres": len(meta["nd"]), "updates": nu,
           "profile_evaluations": B.profile_evaluations(arrays, meta), "ms": t_full * 1e3,
           "updates_per_s": nu / t_full, "tile": {"fpl": tile[0], "threads": tile[1]}, "autotune_ms": times,
           "sample_512": {"persistent_ms": t_sub * 1e3, "plain_ms": t_plain * 1e3, "speedup": t_plain / t_sub,
                          "flux_identical": same}}
    log(f"[bench] batch1d {t_full * 1e3:.1f} ms; 512-sample plain/persistent {t_plain / t_sub:.0f}x")
    del arrays, sa
    return out


def nlte_leg(ctx, args):
    """The 3D NLTE ALI workload (60 x 60 x 120 box, 21-level atom with 80 lines and
    20 continua, 16 rays x 120 frequencies per line): FP32 sweeps with Kahan-summed
    rates, FP64 rate matrices and batched LU on the device, the reference's
    automatic Ng trigger, iterated to max |dn/n| < 1e-4 from LTE.  With --nlte-fp64
    the same iteration with the plain FP64 sweep beside it, run for the same number
    of iterations, and the departure coefficients' largest relative difference
    between the two after them (runs stopped at different iterations would differ
    by the slow modes' remaining drift, several times the 1e-4 step criterion when
    they contract by ~0.95 an iteration, not by the sweep's precision); the FP64
    run's own first iteration under 1e-4 is reported beside it."""
    import numpy as np
    import torch
    from dispatch_amd.ali import AliConfig, NlteProblem
    cfg = AliConfig(max_iter=args.nlte_max_iter, graph=True)
    t0 = time.perf_counter()
    p = NlteProblem(ctx, cfg)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    n = p.solve(log=log)
    its = [dt for _, dt in p.history]
    out = {"workload": "3d_nlte_ali", "box": [cfg.nx, cfg.ny, cfg.nz], "levels": p.atom.nk,
           "lines": p.atom.nline, "continua": len(p.atom.ti) - p.atom.nline, "rays": len(p.vec),
           "frequencies_per_line": cfg.nq, "setup_s": setup, "iterations": n,
           "converged": bool(abs(p.history[-1][0]) < cfg.conv), "iteration_bar": 40,
           "s_per_iteration": float(np.median(its)), "first_iteration_s": its[0],
           "maxerr_last": p.history[-1][0], "ng": "automatic (ng_step 0)",
           "ng_fired": [i + 1 for i, h in enumerate(p.ng_history) if h],
           "line_updates_per_iteration": p.ncell * p.atom.nline * cfg.nq * len(p.vec)}
    out["line_updates_per_s"] = out["line_updates_per_iteration"] / out["s_per_iteration"]
    log(f"[bench] nlte: {n} iterations, {out['s_per_iteration']:.2f} s each")
    if args.nlte_fp64:
        b32 = p.b.clone()
        del p
        torch.cuda.empty_cache()
        q = NlteProblem(ctx, cfg, sweep_dtype=torch.float64)
        q.solve(log=log, iterations=n)
        first = next((i + 1 for i, (e, _) in enumerate(q.history) if abs(e) < cfg.conv), None)
        out["fp64_sweeps"] = {"iterations_run": n, "first_iteration_below_1e-4": first,
                              "s_per_iteration": float(np.median([dt for _, dt in q.history])),
                              "b_max_rel_diff_fp32_vs_fp64": float((b32 / q.b - 1).abs().max()), "tolerance": 1e-4}
        log(f"[bench] nlte fp64: {n} iterations, b differs by {out['fp64_sweeps']['b_max_rel_diff_fp32_vs_fp64']:.2e}")
        del q, b32
    else:
        del p
    torch.cuda.empty_cache()
    return out


VOIGT_PROFILE = os.path.join(ROOT, "profiles", "line_opacity_equi_box120x120x230_nf1024_3lines.md")


def voigt_leg(ctx, box):
    """The cell x frequency Voigt summation of the box's three lines (QLG-bot
    profiles, add_opacity_equi through add_linelist_opacities) over every cell and
    frequency of the 120 x 120 x 230 x 1024 box: one species with the three lines
    (5001, 5002, 5003 A; masses 56, 40, 24 u; centre line-to-continuum ratios 10,
    10^2.5, 10^4; damping a = 0.02, 0.03, 0.05), densities proportional to chi500 so
    the ratios hold at every depth, the continuum chi500 (0.7 + 0.3 (lam/lam0)^3)."""
    import math
    import numpy as np
    import torch
    from dispatch_amd import synth
    from dispatch_amd.constants import BK, CC, UU
    dev = "cuda"
    nf, ncell = box.nf, box.nx * box.ny * box.nz
    lam = torch.as_tensor(box.lam, dtype=torch.float64, device=dev)
    llam = torch.log(lam * 1e8).contiguous()                       # ln(lambda / A)
    T = box.T.reshape(-1).double()
    c5 = box.chi500.reshape(-1).double()
    cont = (c5[None, :] * (0.7 + 0.3 * (lam[:, None] / (5000e-8)) ** 3)).float().contiguous()
    all_eta = torch.zeros_like(cont)
    vmic = 1e5
    lines = [(5001.0, 1e1, 0.02, 56.0), (5002.0, 10 ** 2.5, 0.03, 40.0), (5003.0, 1e4, 0.05, 24.0)]
    isos = []
    # one species, one isotope holding the three lines: n inv_df Rij = ratio * cont / voigt_peak
    awgt = 40.0
    inv_df = CC / torch.sqrt(2 * BK * T / (UU * awgt) + vmic ** 2)
    d = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    vth_c = [math.sqrt(2 * BK * 5800.0 / (UU * m) + vmic ** 2) / CC for _, _, _, m in lines]
    ll = dict(llam0=d([math.log(l0) for l0, _, _, _ in lines], torch.float64),
              lower_ev=d([0.0] * 3, torch.float32),
              gpref=d([math.log(r / 0.5641895835) for _, r, _, _ in lines], torch.float32),
              grad=d([a * v for (_, _, a, _), v in zip(lines, vth_c)], torch.float32),
              gw1=d([0.0] * 3, torch.float32), gw2=d([0.0] * 3, torch.float32), gstk=d([0.0] * 3, torch.float32),
              unsold=d([1] * 3, torch.int32), lev_i=d([0] * 3, torch.int32), lev_j=d([0] * 3, torch.int32))
    isos.append((ll, awgt, 1.0))
    species = [dict(isotopes=isos, n=(c5 / inv_df).contiguous(), part=torch.zeros_like(c5), scale=1.0,
                    scat_group=0)]
    vz = box.vz.reshape(-1).double()
    cells = dict(temp=T.float().contiguous(), vmic=torch.full((ncell,), vmic / 1e5, dtype=torch.float32, device=dev),
                 vproj=(-vz / CC).float().contiguous(), ekt=(11604.518 / T).float().contiguous(),
                 ne=torch.full((ncell,), 1e13, dtype=torch.float32, device=dev),
                 tothi=torch.full((ncell,), 1e16, dtype=torch.float32, device=dev),
                 spe=torch.full((ncell,), 1e13 * 5800 ** (1 / 6), dtype=torch.float32, device=dev))
    all_chi = cont.clone()
    ctx.add_linelist_opacities(species, llam, all_chi, all_eta, cells)      # warm-up (and scratch)
    torch.cuda.synchronize()
    ts = []
    for _ in range(3):
        all_chi.copy_(cont)
        all_eta.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ctx.add_linelist_opacities(species, llam, all_chi, all_eta, cells)
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    ratio = (all_chi / cont - 1.0)
    pairs = 3 * nf * ncell
    out = {"workload": "voigt_sum_3_lines", "cells": ncell, "frequencies": nf, "lines": 3,
           "line_freq_cell_pairs": pairs, "ms": min(ts), "pairs_per_s": pairs / (min(ts) * 1e-3),
           "centre_ratio_max": float(ratio.max()),
           "deposited_fraction": float((all_eta > 0).float().mean())}
    if os.path.exists(VOIGT_PROFILE):
        for line in open(VOIGT_PROFILE):
            if line.startswith("summary:"):
                out["profile"] = line.split(":", 1)[1].strip() + f" (profiles/{os.path.basename(VOIGT_PROFILE)})"
    log(f"[bench] voigt sum {min(ts):.1f} ms over {pairs:.3g} line-frequency-cell pairs")
    del all_chi, all_eta, cont, ratio
    torch.cuda.empty_cache()
    return out


def main():
    # the oracle's OpenMP pool is sized when its library loads: fix the thread count
    # before anything is imported
    os.environ.setdefault("OMP_NUM_THREADS", str(host_cores()))
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the sweep shards by frequency window)", file=sys.stderr)
        sys.exit(2)
    import torch
    from dispatch_amd import m3drt, synth
    torch.cuda.init()
    ctx = m3drt.Context(0)
    if args.calib_copy:
        a = torch.empty(1 << 30, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    t0 = time.time()
    box = synth.make_box(nf=args.nf, nz=args.nz)
    rays, weights = synth.gauss_rays(6)
    cu = [synth.chi500_upwind(box, v) for v in rays]
    log(f"[bench] box ready in {time.time() - t0:.1f}s")
    ncell = box.nx * box.ny * box.nz
    updates = ncell * box.nf * len(rays)

    res = {
        "metric": METRIC, "value": None, "unit": UNIT, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3d_lte_formal_solve", "box": [box.nx, box.ny, box.nz],
                   "slab_planes": 23, "rays": len(rays), "frequencies": box.nf,
                   "updates_per_step": updates, "solver": "short_char2 (linear2)",
                   "layout": "native (x, y, f, z)"},
    }
    s32 = None
    if not args.no_sweep:
        s32 = sweep_legs(ctx, args, box, rays, weights, cu, res)
    if not args.no_adjoint:
        if s32 is not None:
            s32.free()
            torch.cuda.empty_cache()
        res["adjoint"] = adjoint_leg(ctx, args)
        torch.cuda.empty_cache()
    if not args.no_batch1d:
        res["batch1d"] = batch1d_leg(ctx, args)
        torch.cuda.empty_cache()
    if not args.no_output_pass:
        # the output pass on the same box: rotation to the 12 upward rays and the four
        # column solvers over one window of 256 frequencies, the emergent fluxes
        from dispatch_amd import output_pass
        t0 = time.time()
        res["output_pass"] = output_pass.run(nf=256, windows=1)
        log(f"[bench] output pass in {time.time() - t0:.1f}s: {res['output_pass']['ms_per_ray']}")
        torch.cuda.empty_cache()
    if not args.no_voigt:
        res["voigt"] = voigt_leg(ctx, box)
    if not args.no_nlte:
        res["nlte"] = nlte_leg(ctx, args)
    if not args.no_cpu:
        chi32, S32 = synth.build_cubes(box, torch.float32)
        res["cpu_baseline"] = cpu_baseline(box, chi32, S32, rays, weights, cu, args.cpu_seconds)
        del chi32, S32
    ctx.close()
    print(json.dumps(res), flush=True)


def sweep_legs(ctx, args, box, rays, weights, cu, res):
    """The FP32 formal solution (the metric and the roofline) and the plain FP64
    sweep beside it with the FP32 error; returns the FP32 solver when its cubes are
    still held (no FP64 leg), else None."""
    import torch
    ncell = box.nx * box.ny * box.nz
    updates = ncell * box.nf * len(rays)
    s32 = Solver(ctx, box, torch.float32, rays, weights, cu)
    t32 = timed(s32, args.steps, args.warmup)
    ms32 = s32.launch_ms()
    log(f"[bench] fp32 step {t32 * 1e3:.1f} ms  ({updates / t32:.3e} upd/s)")
    res["value"] = updates / t32
    res["ms_per_step"] = t32 * 1e3
    avg = sum(ms32) / len(ms32)
    per_launch_bytes = BYTES_PER_UPDATE["f32"] * ncell * box.nf
    achieved = per_launch_bytes / (avg * 1e-3) / 1e9
    traffic, tkern = measured_traffic(box)
    res["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                       "kernel": "sc2_fast<float> (+ its per-direction dt500 prep)",
                       "avg_launch_ms": avg, "launches": len(ms32),
                       "algorithmic_bytes_per_launch": per_launch_bytes}
    if traffic is not None:
        res["roofline"]["traffic_source"] = f"profiles/{os.path.basename(TRAFFIC_PROFILE)} ({tkern})"

    if not args.no_fp64:
        em32 = s32.emergent.clone()
        J32 = s32.J.clone()
        s32.free()
        torch.cuda.empty_cache()
        s64 = Solver(ctx, box, torch.float64, rays, weights, cu)
        t64 = timed(s64, max(1, args.steps // 2), args.warmup)
        ms64 = s64.launch_ms()
        log(f"[bench] fp64 step {t64 * 1e3:.1f} ms")
        errJ = max_rel_err(J32, s64.J)
        errI = max_rel_err(em32, s64.emergent)
        a64 = sum(ms64) / len(ms64)
        res["fp64"] = {"value": updates / t64, "ms_per_step": t64 * 1e3, "avg_launch_ms": a64,
                       "hbm_GBps": BYTES_PER_UPDATE["f64"] * ncell * box.nf / (a64 * 1e-3) / 1e9,
                       "fp32_speedup": t64 / t32}
        res["accuracy"] = {"J_max_rel_err_fp32_vs_fp64": errJ,
                           "emergent_I_max_rel_err_fp32_vs_fp64": errI, "tolerance": 1e-5}
        del em32, J32
        s64.free()
        torch.cuda.empty_cache()
        return None
    return s32


if __name__ == "__main__":
    main()
