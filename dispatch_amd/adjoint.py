"""Differentiable formal solution: the emergent intensity of a set of rays (and the
mean intensity J) as a torch.autograd.Function of the opacity and source cubes.

Forward: one m3drt_sc_sweep per ray (short_char2 semantics, FP32) writing each
ray's exit plane and, on request, J.  Backward: per ray whose exit plane (or J)
carries a cotangent, the ray's intensity cube is recomputed by the forward sweep
into one scratch cube (nothing per ray is stored between forward and backward:
the checkpoints are the inputs and the ghost planes) and m3drt_sc_adjoint runs the
sweep in reverse, accumulating the cotangents of chi, S and the upstream ghost
planes.  The reference has no adjoint; its forward is rt_patch_t%short_char2
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
    chi500 below the bottom and above the top, the cell sizes."""
    ctx: "m3drt.Context"
    vec: np.ndarray
    wj: np.ndarray
    ds: tuple
    chi500: torch.Tensor
    chi500_up: list
    g500_bot: torch.Tensor
    g500_top: torch.Tensor

    def up(self, r):
        return self.vec[r][2] > 0

    def sweep(self, r, chi, S, ghosts, **out):
        v = self.vec[r]
        up = self.up(r)
        gc, gs, gi = ghosts[0] if up else ghosts[1]
        self.ctx.sc_sweep(chi, S, gc, gs, gi, self.chi500, self.chi500_up[r], self.g500_bot if up else self.g500_top,
                          dir=1 if up else -1, ds=self.ds, vec=v, wj=float(self.wj[r]), **out)

    def adjoint(self, r, chi, S, ghosts, I, gbars, **bars):
        v = self.vec[r]
        up = self.up(r)
        gc, gs, gi = ghosts[0] if up else ghosts[1]
        gcb, gsb, gib = gbars[0] if up else gbars[1]
        self.ctx.sc_adjoint(chi, S, gc, gs, gi, self.chi500, self.chi500_up[r],
                            self.g500_bot if up else self.g500_top, I, dir=1 if up else -1, ds=self.ds, vec=v,
                            wj=float(self.wj[r]), ghost_chibar=gcb, ghost_Sbar=gsb, ghost_Ibar=gib, **bars)


class _FormalSolution(torch.autograd.Function):
    @staticmethod
    def forward(fctx, rs, want_J, chi, S, gc_up, gs_up, gi_up, gc_dn, gs_dn, gi_dn):
        ghosts = ((gc_up, gs_up, gi_up), (gc_dn, gs_dn, gi_dn))
        nz, nf, ny, nx = chi.shape
        exits = torch.zeros((len(rs.vec), nf, ny, nx), dtype=chi.dtype, device=chi.device)
        J = torch.zeros_like(chi) if want_J else None
        for r in range(len(rs.vec)):
            rs.sweep(r, chi, S, ghosts, exit_I=exits[r], J=J)
        fctx.rs = rs
        fctx.save_for_backward(chi, S, gc_up, gs_up, gi_up, gc_dn, gs_dn, gi_dn)
        if want_J:
            return exits, J
        none = torch.zeros((), dtype=chi.dtype, device=chi.device)
        fctx.mark_non_differentiable(none)
        return exits, none

    @staticmethod
    def backward(fctx, exit_bar, J_bar):
        rs = fctx.rs
        chi, S, gc_up, gs_up, gi_up, gc_dn, gs_dn, gi_dn = fctx.saved_tensors
        ghosts = ((gc_up, gs_up, gi_up), (gc_dn, gs_dn, gi_dn))
        if J_bar is not None and J_bar.dim() == 0:
            J_bar = None
        if J_bar is not None:
            J_bar = J_bar.contiguous()
        chibar, Sbar = torch.zeros_like(chi), torch.zeros_like(S)
        gbars = tuple(tuple(torch.zeros_like(t) for t in g) for g in ghosts)
        I = None
        if exit_bar is not None:
            exit_bar = exit_bar.contiguous()
            live = (exit_bar.flatten(1).abs().amax(dim=1) > 0).cpu().numpy()
        else:
            live = np.zeros(len(rs.vec), bool)
        for r in range(len(rs.vec)):
            if not live[r] and J_bar is None:
                continue
            if I is None:
                I = torch.empty_like(chi)
            rs.sweep(r, chi, S, ghosts, I_out=I)
            rs.adjoint(r, chi, S, ghosts, I, gbars, chibar=chibar, Sbar=Sbar, Jbar=J_bar,
                       exit_bar=exit_bar[r] if live[r] else None)
        (gcb_u, gsb_u, gib_u), (gcb_d, gsb_d, gib_d) = gbars
        return None, None, chibar, Sbar, gcb_u, gsb_u, gib_u, gcb_d, gsb_d, gib_d


def formal_solution(rs: RaySet, chi, S, ghost_up, ghost_down, want_J=False):
    """Exit planes of every ray, (nray, nf, ny, nx) (the emergent intensity for the
    upward rays), and J when want_J; differentiable in chi, S and the ghost planes.
    chi, S: FP32 native cubes (nz, nf, ny, nx); ghost_up / ghost_down: (chi, S, I)
    planes (nf, ny, nx) below the bottom and above the top."""
    exits, J = _FormalSolution.apply(rs, want_J, chi, S, *ghost_up, *ghost_down)
    return (exits, J) if want_J else exits
