"""The 3D LTE formal solve on its own workload at full depth: the 120 x 120 x 230
box of dispatch_amd/synth.py (slabs of 23 planes swept as one column), all 24 rays
of the mirrored Gauss set, through the straight-line FP32 sweep kernel the
benchmark times (m3drt_sc_sweep, native layout, J and the exit planes only), on a
subset of 12 of the workload's 1024 frequencies that holds the three line cores
and their wings.  J (summed over the 24 rays with their weights) and the emergent
intensity of the 12 upward rays against the oracle's FP64 short_char2
(SRC/rt_solvers_mod.f90:224-507, with integrate_intensity :3399-3431) on the same
inputs, to 1e-5 relative.
"""
import concurrent.futures as cf

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FSEL = [0, 128, 250, 256, 262, 384, 506, 512, 518, 762, 768, 1023]


@pytest.fixture(scope="module")
def workload():
    from dispatch_amd import synth
    box = synth.make_box(nf=1024, nz=230)
    box.lam = box.lam[FSEL]
    box.nf = len(FSEL)
    chi, S = synth.build_cubes(box, torch.float32)
    rays, weights = synth.gauss_rays(6)
    return box, chi, S, rays, weights


def test_full_depth_formal_solve_matches_fp64_oracle(ctx, workload):
    from dispatch_amd import synth
    box, chi, S, rays, weights = workload
    nz, ny, nx, nf = box.nz, box.ny, box.nx, box.nf
    ghost = synth.ghost_planes(box, chi, torch.float32)
    J = torch.zeros_like(chi)
    nup = int((rays[:, 2] > 0).sum())
    emergent = torch.zeros((nup, nf, ny, nx), dtype=torch.float32, device="cuda")
    cus = [synth.chi500_upwind(box, v) for v in rays]
    iup = 0
    for r, v in enumerate(rays):
        up = v[2] > 0
        gchi, gS, gI = ghost["up" if up else "down"]
        ex = None
        if up:
            ex = emergent[iup]
            iup += 1
        ctx.sc_sweep(chi, S, gchi, gS, gI, box.chi500, cus[r], box.g500_bot if up else box.g500_top,
                     dir=1 if up else -1, ds=box.ds, vec=v, wj=float(weights[r]), J=J, exit_I=ex)
    torch.cuda.synchronize()

    # the oracle: reference layout (nz+2, ny, nx, nf) with the ghost planes, FP64
    chi_r = np.zeros((nz + 2, ny, nx, nf))
    S_r = np.zeros_like(chi_r)
    chi_r[1:-1] = chi.permute(0, 2, 3, 1).double().cpu().numpy()
    S_r[1:-1] = S.permute(0, 2, 3, 1).double().cpu().numpy()
    c500 = box.chi500.cpu().numpy()
    gh = {k: [a.permute(1, 2, 0).double().cpu().numpy() for a in g] for k, g in ghost.items()}
    cu_np = [c.cpu().numpy() for c in cus]
    g500 = {True: box.g500_bot.cpu().numpy(), False: box.g500_top.cpu().numpy()}

    # the work arrays of each sweep direction (shared, read-only), the ghost plane set
    base = {}
    for up in (True, False):
        gz = 0 if up else nz + 1
        c, s = chi_r.copy(), S_r.copy()
        I = np.zeros((nz + 2, ny, nx, nf))
        c[gz], s[gz], I[gz] = gh["up" if up else "down"]
        base[up] = (c, s, I)
    del chi_r, S_r

    def one(r):
        v = rays[r]
        up = bool(v[2] > 0)
        c, s, I = base[up]
        Iout, _, _, _ = O.short_char2(c, s, I, c500, cu_np[r], g500[up], 1 if up else -1, box.ds, v,
                                      get_rates=False, fp64=True)
        return r, Iout[1:-1]

    Jref = np.zeros((nz, ny, nx, nf))
    eref = []
    with cf.ThreadPoolExecutor(8) as pool:
        for r, Iin in pool.map(one, range(len(rays))):
            Jref += weights[r] * Iin
            if rays[r][2] > 0:
                eref.append(Iin[-1])
    Jg = J.permute(0, 2, 3, 1).double().cpu().numpy()
    eg = emergent.permute(0, 2, 3, 1).double().cpu().numpy()
    er = np.stack(eref)
    floor = 1e-6 * np.abs(Jref).max()
    errJ = float(np.max(np.abs(Jg - Jref) / np.maximum(np.abs(Jref), floor)))
    errI = float(np.max(np.abs(eg - er) / np.maximum(np.abs(er), 1e-6 * np.abs(er).max())))
    print(f"full depth 120x120x230, 24 rays, {nf} f: J max rel err {errJ:.2e}, emergent I {errI:.2e}")
    assert errJ < 1e-5
    assert errI < 1e-5
