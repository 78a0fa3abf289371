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


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


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
    p.add_argument("--adjoint-fd", action="store_true",
                   help="check the adjoint's abundance gradients against FP64 central differences")
    p.add_argument("--nf", type=int, default=1024)
    p.add_argument("--nz", type=int, default=230)
    p.add_argument("--cpu-seconds", type=float, default=12.0)
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
    os.environ["OMP_NUM_THREADS"] = str(ncores)
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
    return dict(value=upd / t_used, unit=UNIT, cores=ncores, kind="port",
                sample=f"{done} of {len(rays)} rays x {nfs} of {box.nf} frequencies on the full "
                       f"{nx}x{ny}x{nz} box ({upd:.3g} updates, {t_used:.1f} s); oracle short_char2 "
                       f"FP32 + integrate_intensity, OpenMP over {4 * ncores} frequency windows")


def adjoint_leg(ctx, args):
    """Adjoint / forward time on the abundance-fit workload (96 x 96 x 160, 24 rays,
    2048 frequencies, 12 lines of 4 elements): the forward is the formal solution as
    configured (24 sweeps, J and the emergent intensities); the adjoint is the
    gradient of the flux chi^2 with respect to chi and S (the 12 upward rays, each
    recomputed by the forward sweep and swept back), then to the 4 log abundances."""
    import torch
    from dispatch_amd.fit import FitProblem
    t0 = time.time()
    torch.cuda.reset_peak_memory_stats()
    p = FitProblem(ctx)
    p.make_target()
    log(f"[bench] fit problem ready in {time.time() - t0:.1f}s")
    d0 = [0.0, 0.0, 0.0, 0.0]
    chi, S = p.cubes(d0)
    fwd = []
    for i in range(1 + max(1, args.steps // 2)):
        torch.cuda.synchronize()
        a = time.perf_counter()
        with torch.no_grad():
            p.forward_full(chi, S)
        torch.cuda.synchronize()
        if i:
            fwd.append(time.perf_counter() - a)
    del chi, S
    tm = {}
    for i in range(2):
        loss, g, _ = p.loss_and_grad(d0, timing=tm)
    t_fwd = min(fwd)
    out = {"workload": "abundance_fit", "box": [p.box.nx, p.box.ny, p.box.nz], "rays": len(p.vec),
           "frequencies": p.box.nf, "lines": 12, "elements": 4,
           "forward_ms": t_fwd * 1e3, "adjoint_sweeps_ms": tm["adjoint_s"] * 1e3,
           "abundance_chain_ms": tm["chain_s"] * 1e3,
           "adjoint_over_forward": tm["adjoint_s"] / t_fwd,
           "gradient_over_forward": (tm["adjoint_s"] + tm["chain_s"]) / t_fwd,
           "peak_hbm_GB": torch.cuda.max_memory_allocated() / 1e9,
           "loss": loss, "grad_log_abundance": [float(v) for v in g]}
    log(f"[bench] adjoint {tm['adjoint_s'] * 1e3:.1f} ms vs forward {t_fwd * 1e3:.1f} ms")
    if args.adjoint_fd:
        fd = []
        for e in range(4):
            dp, dm = list(d0), list(d0)
            dp[e] += 1e-3
            dm[e] -= 1e-3
            fd.append((float(p.loss_of_flux(p.flux_fp64(dp))) - float(p.loss_of_flux(p.flux_fp64(dm)))) / 2e-3)
        out["fd_fp64_h1e-3"] = fd
        out["fd_max_rel_err"] = max(abs(float(a) - b) / abs(b) for a, b in zip(g, fd))
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


def main():
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

    s32 = Solver(ctx, box, torch.float32, rays, weights, cu)
    t32 = timed(s32, args.steps, args.warmup)
    ms32 = s32.launch_ms()
    log(f"[bench] fp32 step {t32 * 1e3:.1f} ms  ({updates / t32:.3e} upd/s)")
    res = {
        "metric": METRIC, "value": updates / t32, "unit": UNIT, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t32 * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3d_lte_formal_solve", "box": [box.nx, box.ny, box.nz],
                   "slab_planes": 23, "rays": len(rays), "frequencies": box.nf,
                   "updates_per_step": updates, "solver": "short_char2 (linear2)",
                   "layout": "native (x, y, f, z)"},
    }
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
        chi_src = None
    if not args.no_adjoint:
        if args.no_fp64:
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
    if not args.no_cpu:
        chi32, S32 = synth.build_cubes(box, torch.float32)
        res["cpu_baseline"] = cpu_baseline(box, chi32, S32, rays, weights, cu, args.cpu_seconds)
        del chi32, S32
    ctx.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
