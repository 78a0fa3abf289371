"""Abundance-fit workload on the differentiable formal solution.

A 96 x 96 x 160 box from the 3D LTE generator (dispatch_amd.synth.make_box: grey
T(tau) plus the tapered random field, hydrostatic chi500, Gaussian vertical
velocities), 24 rays (2 Gauss mu x 6 azimuths, mirrored), 2048 frequencies over
5000-5008 A holding 12 synthetic Voigt lines of 4 elements on the grey + lambda^3
continuum; chi = chi500 (continuum + sum_e 10^(d_e) L_e(T)) and S = B_nu(T), with
d_e the offsets of the 4 log abundances from the reference ones.  The observable
is the disk-integrated flux (the mu-weighted mean emergent intensity of the 12
upward rays, averaged over the surface); the loss is chi^2 against a target made
at offsets of +-0.2 dex with Gaussian noise at S/N 300.

Gradients: the adjoint sweep (dispatch_amd.adjoint, m3drt_sc_adjoint) takes the
flux cotangent back to chi and S; the abundance gradient is then
dL/dd_e = ln10 10^d_e sum chi500 L_e chibar (a reduction per element, over the
element cubes chi500 L_e, which are computed once: they do not depend on the
abundances), and the temperature gradient comes from torch autograd through the
opacity of each plane, recomputed.  chi500 (the tau500 step scale) and the velocities are held
fixed.  The FP64 path (FP64 cubes and the plain FP64 sweep) gives the central
finite differences the gradients are checked against.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import m3drt, synth
from .adjoint import RaySet, formal_solution

ELEMENTS = ("e1", "e2", "e3", "e4")
MASS = (56.0, 40.0, 24.0, 48.0)
# (element, centre offset from 5000 A [A], line-to-continuum ratio at the centre, damping a)
LINES = tuple((l % 4, 0.35 + 0.62 * l, 10 ** (0.6 + 0.27 * ((7 * l) % 12)), 0.01 + 0.004 * (l % 5))
              for l in range(12))
TRUE_OFFSETS = (0.2, -0.2, 0.2, -0.2)
SLAB = 16


class FitProblem:
    def __init__(self, ctx, nx=96, ny=96, nz=160, nf=2048, seed=20261206, nphi=6):
        self.ctx = ctx
        self.box = synth.make_box(nx, ny, nz, nf, seed=seed, band_A=8.0)
        vec, wj = synth.gauss_rays(nphi)
        self.vec, self.wj = vec, wj
        box = self.box
        self.rs = RaySet(ctx, vec, wj, box.ds, box.chi500, [synth.chi500_upwind(box, v) for v in vec],
                         box.g500_bot, box.g500_top)
        up = vec[:, 2] > 0
        # flux weights: 2 w_mu mu over the upward half (the J weights of the mirrored set sum to 1)
        self.wf = torch.tensor(np.where(up, 2.0 * wj * vec[:, 2], 0.0), dtype=torch.float64, device="cuda")
        self.lam = torch.as_tensor(box.lam, dtype=torch.float64, device="cuda")
        self.r500 = (box.g500_bot.double() / box.chi500[0].double())

    # -------------------------------------------------------------- opacity
    def slab(self, z0, z1, T=None):
        """(cont, L (4 element line opacities per chi500), S) for planes z0..z1 in
        FP64, native layout (nz, nf, ny, nx); T (z1-z0, ny, nx) may require grad."""
        box = self.box
        lam = self.lam[None, :, None, None]
        T = (box.T[z0:z1] if T is None else T)[:, None]
        vz = box.vz[z0:z1][:, None]
        cont = (0.7 + 0.3 * (lam / 5000e-8) ** 3).expand(z1 - z0, -1, box.ny, box.nx)
        L = [torch.zeros((z1 - z0, box.nf, box.ny, box.nx), dtype=torch.float64, device="cuda") for _ in ELEMENTS]
        for e, off, ratio, a in LINES:
            lc = (5000.0 + off) * 1e-8 * (1.0 + vz / synth.C_LIGHT)
            vth = torch.sqrt(2 * synth.K_BOLTZ * T / (MASS[e] * synth.AMU) + 1e5 ** 2)
            v = (lam - lc) / (lc * vth / synth.C_LIGHT)
            L[e] = L[e] + ratio * (T / 5800.0) ** -3 * synth._voigt(a, v)
        S = synth.planck_occupation(lam, T)
        return cont, L, S

    def chi_S(self, d, z0, z1, T=None):
        cont, L, S = self.slab(z0, z1, T)
        c5 = self.box.chi500[z0:z1].double()[:, None]
        tot = cont
        for e in range(len(ELEMENTS)):
            tot = tot + 10.0 ** d[e] * L[e]
        return c5 * tot, S

    def cubes(self, d, dtype=torch.float32):
        """chi and S for abundance offsets d.  FP32 from the cached element cubes (the
        opacity's T dependence is fixed while only the abundances move); FP64 (the
        finite-difference reference) from the opacity at the current T, each value
        rounded to FP32 first so both see the same inputs."""
        box = self.box
        if dtype == torch.float32:
            self._prepare()
            chi = self._cont.clone()
            for e in range(len(ELEMENTS)):
                chi.add_(self._L[e], alpha=10.0 ** float(d[e]))
            return chi, self._S.clone()
        chi = torch.empty((box.nz, box.nf, box.ny, box.nx), dtype=dtype, device="cuda")
        S = torch.empty_like(chi)
        d = torch.as_tensor(d, dtype=torch.float64, device="cuda")
        for z0 in range(0, box.nz, SLAB):
            z1 = min(box.nz, z0 + SLAB)
            c, s = self.chi_S(d, z0, z1)
            chi[z0:z1] = c.float().to(dtype)
            S[z0:z1] = s.float().to(dtype)
        return chi, S

    def _prepare(self):
        """The FP32 cubes chi500 cont, chi500 L_e (e = 1..4) and S at the current T."""
        if getattr(self, "_cont", None) is not None:
            return
        box = self.box
        shape = (box.nz, box.nf, box.ny, box.nx)
        self._cont = torch.empty(shape, dtype=torch.float32, device="cuda")
        self._L = [torch.empty_like(self._cont) for _ in ELEMENTS]
        self._S = torch.empty_like(self._cont)
        for z0 in range(0, box.nz, SLAB):
            z1 = min(box.nz, z0 + SLAB)
            cont, L, S = self.slab(z0, z1)
            c5 = box.chi500[z0:z1].double()[:, None]
            self._cont[z0:z1] = (c5 * cont).float()
            for e in range(len(ELEMENTS)):
                self._L[e][z0:z1] = (c5 * L[e]).float()
            self._S[z0:z1] = S.float()

    def release(self):
        """Drop the cached cubes (after a change of T)."""
        self._cont = self._L = self._S = None

    def ghosts(self, chi):
        """The LTE plane below the bottom (chi(z=1) chi500_ghost/chi500(1), S = I =
        B_nu(T_ghost); SRC/rt_patch_mod.f90:3352-3373) and zeros above the top."""
        box = self.box
        lam = self.lam[:, None, None]
        cb = (chi[0].double() * self.r500[None]).to(chi.dtype)
        Sb = synth.planck_occupation(lam, box.T_ghost_bot[None]).float().to(chi.dtype)
        z = torch.zeros_like(cb)
        return (cb.contiguous(), Sb.contiguous(), Sb.clone()), (z, z.clone(), z.clone())

    # -------------------------------------------------------------- observable
    def flux(self, exits):
        return torch.einsum("r,rf->f", self.wf.to(exits.dtype), exits.mean(dim=(2, 3)))

    def forward_full(self, chi, S):
        """The formal solution as configured: 24 rays in one launch each, J and the
        exit planes (no checkpoints: nothing is kept for a backward pass)."""
        gu, gd = self.ghosts(chi)
        box = self.box
        exits = torch.zeros((len(self.vec), box.nf, box.ny, box.nx), dtype=chi.dtype, device="cuda")
        J = torch.zeros_like(chi)
        for r in range(len(self.vec)):
            self.rs.sweep(r, chi, S, (gu, gd), exit_I=exits[r], J=J)
        return self.flux(exits), J

    def up_rays(self):
        return [r for r, v in enumerate(self.vec) if v[2] > 0]

    def forward_flux(self, chi, S):
        """The forward the flux loss needs: the upward rays' exit planes, one launch
        per ray, no J and no checkpoints."""
        gu, gd = self.ghosts(chi)
        box = self.box
        exits = torch.zeros((len(self.vec), box.nf, box.ny, box.nx), dtype=chi.dtype, device="cuda")
        for r in self.up_rays():
            self.rs.sweep(r, chi, S, (gu, gd), exit_I=exits[r])
        return self.flux(exits)

    def flux_fp64(self, d):
        """Flux of the plain FP64 sweep (upward rays only: the downward ones do not
        reach the surface) on FP64 cubes of the same FP32 values."""
        chi, S = self.cubes(d, torch.float64)
        gu, _ = self.ghosts(chi)
        box = self.box
        F = torch.zeros(box.nf, dtype=torch.float64, device="cuda")
        ex = torch.empty((box.nf, box.ny, box.nx), dtype=torch.float64, device="cuda")
        for r, v in enumerate(self.vec):
            if v[2] <= 0:
                continue
            self.ctx.sc_sweep(chi, S, *gu, box.chi500, self.rs.chi500_up[r], box.g500_bot, dir=1, ds=box.ds,
                              vec=v, exit_I=ex)
            F += self.wf[r] * ex.mean(dim=(1, 2))
        del chi, S
        return F

    def make_target(self, offsets=TRUE_OFFSETS, snr=300.0, seed=7):
        chi, S = self.cubes(offsets)
        with torch.no_grad():
            F, _ = self.forward_full(chi, S)
        del chi, S
        g = torch.Generator(device="cuda").manual_seed(seed)
        self.sigma = F.double() / snr
        self.target = F.double() + self.sigma * torch.randn(F.shape, generator=g, dtype=torch.float64, device="cuda")

    def loss_of_flux(self, F):
        return (((F.double() - self.target) / self.sigma) ** 2).sum()

    # -------------------------------------------------------------- gradients
    def loss_and_grad(self, d, want_T=False, timing=None):
        """chi^2 and its gradient with respect to the 4 abundance offsets (and per-cell
        T when want_T), from the FP32 forward and the adjoint sweep."""
        d = torch.as_tensor(d, dtype=torch.float64, device="cuda")
        chi, S = self.cubes(d)
        chi.requires_grad_()
        S.requires_grad_()
        gu, gd = self.ghosts(chi.detach())
        gu = tuple(t.requires_grad_() for t in gu)
        t0 = _now()
        exits = formal_solution(self.rs, chi, S, gu, gd, rays=self.up_rays())
        loss = self.loss_of_flux(self.flux(exits))
        t1 = _now()
        loss.backward()
        t2 = _now()
        chibar = chi.grad
        chibar[0] += (gu[0].grad.double() * self.r500[None]).float()  # the ghost's chi follows chi(z=1)
        # dL/dd_e = ln10 10^d_e <chi500 L_e, chibar>: one dot product per element and
        # slab of planes (FP32 within a slab, FP64 across)
        gd_ = torch.zeros(len(ELEMENTS), dtype=torch.float64, device="cuda")
        box = self.box
        with torch.no_grad():
            for e in range(len(ELEMENTS)):
                acc = torch.zeros((), dtype=torch.float64, device="cuda")
                for z0 in range(0, box.nz, SLAB):
                    z1 = min(box.nz, z0 + SLAB)
                    acc += torch.dot(self._L[e][z0:z1].reshape(-1), chibar[z0:z1].reshape(-1)).double()
                gd_[e] = math.log(10.0) * 10.0 ** float(d[e]) * acc
        gT = None
        if want_T:
            # one plane at a time: autograd keeps every intermediate of the profiles
            gT = torch.zeros_like(box.T)
            for z in range(box.nz):
                T = box.T[z:z + 1].clone().requires_grad_()
                c, s = self.chi_S(d, z, z + 1, T)
                ((c * chibar[z:z + 1].double()).sum() + (s * S.grad[z:z + 1].double()).sum()).backward()
                gT[z] = T.grad[0]
        t3 = _now()
        if want_T:
            # the ghost plane's S = I = B_nu(T_ghost), T_ghost = 2 T(1) - T(2)
            Tg = box.T_ghost_bot.clone().requires_grad_()
            B = synth.planck_occupation(self.lam[:, None, None], Tg[None])
            (B * (gu[1].grad.double() + gu[2].grad.double())).sum().backward()
            gT[0] += 2 * Tg.grad
            gT[1] -= Tg.grad
        if timing is not None:
            timing.update(forward_s=t1 - t0, adjoint_s=t2 - t1, chain_s=t3 - t2)
        return float(loss.detach()), gd_, gT


    # -------------------------------------------------------------- fit
    def lbfgs(self, steps=50, d0=(0.0, 0.0, 0.0, 0.0), log=None):
        """Fit the 4 log-abundance offsets by L-BFGS (torch.optim.LBFGS, strong
        Wolfe line search, history 10) from d0, `steps` iterations; every loss and
        gradient from the FP32 forward and the adjoint.  Returns the trajectory."""
        d = torch.tensor(d0, dtype=torch.float64, device="cuda", requires_grad=True)
        opt = torch.optim.LBFGS([d], lr=1.0, max_iter=1, history_size=10, line_search_fn="strong_wolfe",
                                tolerance_grad=0.0, tolerance_change=0.0)
        evals = []

        def closure():
            opt.zero_grad()
            loss, g, _ = self.loss_and_grad(d.detach())
            d.grad = g.clone()
            evals.append(loss)
            return torch.tensor(loss, dtype=torch.float64, device="cuda")

        traj = []
        t0 = time.perf_counter()
        for it in range(steps):
            opt.step(closure)
            traj.append(dict(step=it + 1, loss=evals[-1], d=[float(v) for v in d.detach()], evaluations=len(evals),
                             s=time.perf_counter() - t0))
            if log:
                log(f"[fit] step {it + 1}: loss {evals[-1]:.6g} d {traj[-1]['d']}")
        return traj


def _now():
    torch.cuda.synchronize()
    return time.perf_counter()


def main():
    """python -m dispatch_amd.fit [--steps 50]: the abundance fit of the 96 x 96 x 160
    x 2048 workload from zero offsets against the target at +-0.2 dex (S/N 300);
    prints one JSON line with the loss trajectory, the fitted offsets, the time per
    step and the peak HBM."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=50)
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    ctx = m3drt.Context(0)
    torch.cuda.reset_peak_memory_stats()
    p = FitProblem(ctx)
    p.make_target()
    loss0, _, _ = p.loss_and_grad([0.0] * 4)
    traj = p.lbfgs(args.steps, log=log)
    out = dict(workload="abundance_fit_lbfgs", box=[p.box.nx, p.box.ny, p.box.nz], frequencies=p.box.nf,
               rays=len(p.vec), true_offsets=list(TRUE_OFFSETS), loss_start=loss0, steps=args.steps,
               loss_end=traj[-1]["loss"], d_end=traj[-1]["d"], evaluations=traj[-1]["evaluations"],
               s_total=traj[-1]["s"], s_per_evaluation=traj[-1]["s"] / traj[-1]["evaluations"],
               peak_hbm_GB=torch.cuda.max_memory_allocated() / 1e9,
               loss_trajectory=[t["loss"] for t in traj], d_trajectory=[t["d"] for t in traj])
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
