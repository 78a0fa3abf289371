"""Differentiable formal solution: the emergent intensity of a set of rays (and the
mean intensity J) as a torch.autograd.Function of the opacity and source cubes.

Forward: each ray is swept patch by patch, as DISPATCH sweeps a column of patches
(`patch` planes each, 23 by default): one m3drt_sc_sweep per patch whose upstream
ghost plane is the previous patch's last plane (chi, S) with the intensity that
patch left on its exit plane (set_upstream_boundary / set_downstream_boundary,
SRC/rt_patch_mod.f90:3270-3300, 3471-3511).  The exit planes of the patches are the
checkpoints: one (nf, ny, nx) plane per patch boundary and ray, kept for the
backward pass.  The patch sweeps give the single-launch sweep's results bit for bit
(the ghost plane carries exactly the values the upwind plane holds).

Backward: per ray whose exit plane (or J) carries a cotangent, the patches are
taken in reverse sweep order: the patch's intensity is recomputed from its
checkpoint into one patch-sized scratch, and m3drt_sc_adjoint runs the patch in
reverse, accumulating the cotangents of its chi and S; the cotangent it returns
for its ghost plane is that of the previous patch's exit plane (I) and last plane
(chi, S), and is handed on.  Nothing larger than a patch is recomputed or stored.
The reference has no adjoint; its forward is rt_patch_t%short_char2
(SRC/rt_solvers_mod.f90:224-507) with integrate_intensity
(SRC/rt_patch_mod.f90:3399-3431).

chi500 (the tau500 step scale) and the directions are held fixed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import m3drt


@dataclass
class RaySet:
    """The directions of one formal solution over a box in the native layout
    (nz, nf, ny, nx): unit vectors (nray, 3), J weights (nray), chi500 (nz, ny, nx),
    chi500 at each ray's upwind points (list of (nz, ny, nx)), the ghost planes'
    chi500 below the bottom and above the top, the cell sizes; `patch` planes per
    patch of the checkpointed sweep."""
    ctx: "m3drt.Context"
    vec: np.ndarray
    wj: np.ndarray
    ds: tuple
    chi500: torch.Tensor
    chi500_up: list
    g500_bot: torch.Tensor
    g500_top: torch.Tensor
    patch: int = 23

    def up(self, r):
        return self.vec[r][2] > 0

    def sweep(self, r, chi, S, ghosts, **out):
        """The whole column in one launch."""
        v = self.vec[r]
        up = self.up(r)
        gc, gs, gi = ghosts[0] if up else ghosts[1]
        self.ctx.sc_sweep(chi, S, gc, gs, gi, self.chi500, self.chi500_up[r], self.g500_bot if up else self.g500_top,
                          dir=1 if up else -1, ds=self.ds, vec=v, wj=float(self.wj[r]), **out)

    # ---------------------------------------------------------------- patches
    def patches(self, r, nz):
        """(z0, z1) of every patch in sweep order."""
        b = [(z0, min(nz, z0 + self.patch)) for z0 in range(0, nz, self.patch)]
        return b if self.up(r) else b[::-1]

    def _ghost_of(self, r, k, bounds, chi, S, ghosts, ckpt):
        """The upstream ghost (chi, S, I) and its chi500 of patch k (sweep order)."""
        up = self.up(r)
        if k == 0:
            return (ghosts[0] if up else ghosts[1]), (self.g500_bot if up else self.g500_top)
        z0, z1 = bounds[k - 1]
        zl = z1 - 1 if up else z0          # the previous patch's last plane in sweep order
        return (chi[zl], S[zl], ckpt[k - 1]), self.chi500[zl]

    def sweep_patch(self, r, k, bounds, chi, S, ghosts, ckpt, J=None, **out):
        """Patch k (sweep order) of ray r: J (a whole cube) accumulates over the
        patch's planes; the other outputs (exit_I, I_out) are the patch's own."""
        (gc, gs, gi), g500 = self._ghost_of(r, k, bounds, chi, S, ghosts, ckpt)
        z0, z1 = bounds[k]
        if J is not None:
            out["J"] = J[z0:z1]
        self.ctx.sc_sweep(chi[z0:z1], S[z0:z1], gc, gs, gi, self.chi500[z0:z1], self.chi500_up[r][z0:z1], g500,
                          dir=1 if self.up(r) else -1, ds=self.ds, vec=self.vec[r], wj=float(self.wj[r]), **out)

    def sweep_checkpointed(self, r, chi, S, ghosts, ckpt, exit_I, J=None):
        """The ray patch by patch: ckpt (npatch - 1, nf, ny, nx) receives the exit
        plane of every patch but the last, whose exit plane goes to exit_I."""
        bounds = self.patches(r, chi.shape[0])
        for k in range(len(bounds)):
            ex = ckpt[k] if k + 1 < len(bounds) else exit_I
            self.sweep_patch(r, k, bounds, chi, S, ghosts, ckpt, exit_I=ex, J=J)

    def adjoint_checkpointed(self, r, chi, S, ghosts, ckpt, scratch, gbars, chibar, Sbar, Jbar=None,
                             exit_bar=None, work=None):
        """The reverse of sweep_checkpointed: patch by patch from the exit, each
        patch's intensity recomputed from its checkpoint into scratch (>= patch
        planes).  gbars: the true ghost planes' cotangents (accumulated); work: three
        (nf, ny, nx) planes for the patch-boundary cotangents."""
        bounds = self.patches(r, chi.shape[0])
        up = self.up(r)
        gcb, gsb, gib = work
        carry = exit_bar
        for k in range(len(bounds) - 1, -1, -1):
            z0, z1 = bounds[k]
            n = z1 - z0
            I = scratch[:n]
            self.sweep_patch(r, k, bounds, chi, S, ghosts, ckpt, I_out=I)
            (gc, gs, gi), g500 = self._ghost_of(r, k, bounds, chi, S, ghosts, ckpt)
            if k == 0:
                tb = gbars[0] if up else gbars[1]
                ghost_bars = dict(ghost_chibar=tb[0], ghost_Sbar=tb[1], ghost_Ibar=tb[2])
            else:
                for t in work:
                    t.zero_()
                ghost_bars = dict(ghost_chibar=gcb, ghost_Sbar=gsb, ghost_Ibar=gib)
            self.ctx.sc_adjoint(chi[z0:z1], S[z0:z1], gc, gs, gi, self.chi500[z0:z1], self.chi500_up[r][z0:z1],
                                g500, I, dir=1 if up else -1, ds=self.ds, vec=self.vec[r], wj=float(self.wj[r]),
                                chibar=chibar[z0:z1], Sbar=Sbar[z0:z1],
                                Jbar=None if Jbar is None else Jbar[z0:z1], exit_bar=carry, **ghost_bars)
            if k > 0:
                pz0, pz1 = bounds[k - 1]
                zl = pz1 - 1 if up else pz0
                chibar[zl] += gcb
                Sbar[zl] += gsb
                carry = gib.clone()


class _FormalSolution(torch.autograd.Function):
    @staticmethod
    def forward(fctx, rs, want_J, rays, chi, S, gc_up, gs_up, gi_up, gc_dn, gs_dn, gi_dn):
        ghosts = ((gc_up, gs_up, gi_up), (gc_dn, gs_dn, gi_dn))
        nz, nf, ny, nx = chi.shape
        exits = torch.zeros((len(rs.vec), nf, ny, nx), dtype=chi.dtype, device=chi.device)
        J = torch.zeros_like(chi) if want_J else None
        npatch = (nz + rs.patch - 1) // rs.patch
        ckpt = {}
        for r in rays:
            ckpt[r] = torch.empty((max(npatch - 1, 1), nf, ny, nx), dtype=chi.dtype, device=chi.device)
            rs.sweep_checkpointed(r, chi, S, ghosts, ckpt[r], exits[r], J=J)
        fctx.rs, fctx.rays, fctx.ckpt = rs, rays, ckpt
        fctx.save_for_backward(chi, S, gc_up, gs_up, gi_up, gc_dn, gs_dn, gi_dn)
        if want_J:
            return exits, J
        none = torch.zeros((), dtype=chi.dtype, device=chi.device)
        fctx.mark_non_differentiable(none)
        return exits, none

    @staticmethod
    def backward(fctx, exit_bar, J_bar):
        rs, rays, ckpt = fctx.rs, fctx.rays, fctx.ckpt
        chi, S, gc_up, gs_up, gi_up, gc_dn, gs_dn, gi_dn = fctx.saved_tensors
        ghosts = ((gc_up, gs_up, gi_up), (gc_dn, gs_dn, gi_dn))
        if J_bar is not None and J_bar.dim() == 0:
            J_bar = None
        if J_bar is not None:
            J_bar = J_bar.contiguous()
        chibar, Sbar = torch.zeros_like(chi), torch.zeros_like(S)
        gbars = tuple(tuple(torch.zeros_like(t) for t in g) for g in ghosts)
        if exit_bar is not None:
            exit_bar = exit_bar.contiguous()
            live = (exit_bar.flatten(1).abs().amax(dim=1) > 0).cpu().numpy()
        else:
            live = np.zeros(len(rs.vec), bool)
        scratch = work = None
        for r in rays:
            if not live[r] and J_bar is None:
                continue
            if scratch is None:
                scratch = torch.empty((min(rs.patch, chi.shape[0]),) + tuple(chi.shape[1:]), dtype=chi.dtype,
                                      device=chi.device)
                work = tuple(torch.zeros_like(gc_up) for _ in range(3))
            rs.adjoint_checkpointed(r, chi, S, ghosts, ckpt[r], scratch, gbars, chibar, Sbar, Jbar=J_bar,
                                    exit_bar=exit_bar[r] if live[r] else None, work=work)
        fctx.ckpt = None
        (gcb_u, gsb_u, gib_u), (gcb_d, gsb_d, gib_d) = gbars
        return None, None, None, chibar, Sbar, gcb_u, gsb_u, gib_u, gcb_d, gsb_d, gib_d


def formal_solution(rs: RaySet, chi, S, ghost_up, ghost_down, want_J=False, rays=None):
    """Exit planes of every ray, (nray, nf, ny, nx) (the emergent intensity for the
    upward rays), and J when want_J; differentiable in chi, S and the ghost planes.
    chi, S: FP32 native cubes (nz, nf, ny, nx); ghost_up / ghost_down: (chi, S, I)
    planes (nf, ny, nx) below the bottom and above the top.  rays: the indices of
    the rays to sweep (default all; the others' exit planes stay zero and they add
    nothing to J)."""
    rays = list(range(len(rs.vec))) if rays is None else [int(r) for r in rays]
    exits, J = _FormalSolution.apply(rs, want_J, rays, chi, S, *ghost_up, *ghost_down)
    return (exits, J) if want_J else exits
