"""GPU parity of the adjoint sweep (m3drt_sc_adjoint) against the oracle's FP64
reverse of short_char2 (oracle/m3d_adjoint.c, itself pinned by central finite
differences of the FP64 forward).

The kernel differentiates the FP32 forward (hardware exponential, FP32 upwind
sums) and takes the forward intensity from the FP32 sweep; the oracle differentiates
the FP64 forward with exact weights.  Bar: every cotangent array within 2e-5 of its
largest magnitude (the FP32 rounding of the forward and of the reverse sums).

Two kernel forms: the strip form (a thread per column and strip of rows, where a strip
length of 4, 6, 8 or 12 divides ny and the strips fit one workgroup) and the
cell-owner form (forced with M3DRT_ADJ_CELLS, the only one for the other planes);
each case runs through both.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from sweep_cases import interior, make_case, plane_to_native, to_native

pytestmark = pytest.mark.gpu
TOL = 2e-5


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def layout_fns(layout, gz):
    if layout == "native":
        return (lambda a: dev(to_native(a)), lambda a: dev(plane_to_native(a)),
                lambda t: np.moveaxis(t.double().cpu().numpy(), 1, 3), lambda t: np.moveaxis(t.double().cpu().numpy(), 0, 2))
    if layout == "fmajor":
        return (lambda a: dev(np.moveaxis(a, 3, 0)), lambda a: dev(plane_to_native(a)),
                lambda t: np.moveaxis(t.double().cpu().numpy(), 0, 3), lambda t: np.moveaxis(t.double().cpu().numpy(), 0, 2))
    return dev, dev, lambda t: t.double().cpu().numpy(), lambda t: t.double().cpu().numpy()


def close(got, ref, what):
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < TOL, f"{what}: {err:.3e}"


CASES = [dict(seed=1, vec=(0.3, -0.4, 0.866)), dict(seed=2, vec=(-0.5, 0.2, -0.84)),
         dict(seed=3, vec=(0.61, 0.05, -0.79), top_ghost_zero=True),
         dict(seed=4, vec=(-0.2, -0.7, 0.69), nx=9, ny=8, nz=6, nf=2),
         dict(seed=5, vec=(0.9, 0.2, 0.39), nx=3, ny=5, nz=4),
         dict(seed=6, vec=(0.3, -0.4, 0.866), nx=40, ny=30, nz=4, nf=3),
         dict(seed=7, vec=(-0.45, 0.35, -0.82), nx=96, ny=96, nz=3, nf=2),
         dict(seed=8, vec=(0.3, -0.4, 0.866), nx=20, ny=32, nz=3, nf=2),
         dict(seed=9, vec=(0.55, 0.5, 0.67), nx=10, ny=12, nz=4, nf=2)]


@pytest.mark.parametrize("form", ["strip", "cells"])
@pytest.mark.parametrize("layout", ["native", "reference", "fmajor"])
@pytest.mark.parametrize("kw", CASES)
def test_adjoint_matches_oracle(ctx, kw, layout, form, monkeypatch):
    if form == "cells":
        monkeypatch.setenv("M3DRT_ADJ_CELLS", "1")
    kw = dict(kw)
    tgz = kw.pop("top_ghost_zero", False)
    case = make_case(nf=kw.pop("nf", 3), nx=kw.pop("nx", 7), ny=kw.pop("ny", 6), nz=kw.pop("nz", 5),
                     top_ghost_zero=tgz, **kw)
    print(f"adjoint {form} {layout} nf={case['nf']} plane={case['nx']}x{case['ny']} nz={case['nz']} vec={case['vec']}",
          flush=True)
    chi, S = case["chi"].astype(np.float64), case["S"].astype(np.float64)
    I64 = O.short_char2(chi, S, case["I"].astype(np.float64), case["chi500"], case["chi500_up"], case["g500"],
                        case["dir"], case["ds"], case["vec"], get_rates=False, fp64=True)[0]
    gz = case["gz"]
    rng = np.random.default_rng(kw["seed"])
    Jb = rng.uniform(-1, 1, I64[1:-1].shape)
    Ib = rng.uniform(-1, 1, I64[1:-1].shape)
    eb = rng.uniform(-1, 1, I64[gz].shape)
    wj = 0.375
    # the oracle's direct cotangent of every plane: Ibar + wj Jbar, the exit plane's added
    w = np.zeros_like(chi)
    w[1:-1] = Ib + wj * Jb
    last = case["nz"] if case["dir"] > 0 else 1
    w[last] += eb
    rIb, rcb, rsb = O.short_char2_adjoint(chi, S, I64, case["chi500"], case["chi500_up"], case["g500"], w,
                                          case["dir"], case["ds"], case["vec"])
    cub, pla, back, pback = layout_fns(layout, gz)
    c, s = cub(interior(case["chi"])), cub(interior(case["S"]))
    gc, gs, gi = pla(case["chi"][gz]), pla(case["S"][gz]), pla(case["I"][gz])
    c500, cu, g5 = dev(case["chi500"]), dev(case["chi500_up"]), dev(case["g500"])
    Iout = torch.zeros_like(c)
    ctx.sc_sweep(c, s, gc, gs, gi, c500, cu, g5, dir=case["dir"], ds=case["ds"], vec=case["vec"], layout=layout,
                 I_out=Iout)
    cb, sb = torch.zeros_like(c), torch.zeros_like(c)
    gcb, gsb, gib = torch.zeros_like(gc), torch.zeros_like(gc), torch.zeros_like(gc)
    ctx.sc_adjoint(c, s, gc, gs, gi, c500, cu, g5, Iout, dir=case["dir"], ds=case["ds"], vec=case["vec"],
                   layout=layout, chibar=cb, Sbar=sb, wj=wj, Jbar=cub(Jb), Ibar=cub(Ib), exit_bar=pla(eb),
                   ghost_chibar=gcb, ghost_Sbar=gsb, ghost_Ibar=gib)
    torch.cuda.synchronize()
    close(back(cb), rcb[1:-1], "chibar")
    close(back(sb), rsb[1:-1], "Sbar")
    close(pback(gcb), rcb[gz], "ghost chibar")
    close(pback(gsb), rsb[gz], "ghost Sbar")
    close(pback(gib), rIb[gz], "ghost Ibar")


def test_adjoint_accumulates_and_rejects(ctx):
    """A second call adds the same cotangents again; unsupported forms are errors."""
    from dispatch_amd import m3drt
    case = make_case(nf=2, nx=8, ny=6, nz=4, seed=9)
    gz = case["gz"]
    cub, pla, back, _ = layout_fns("native", gz)
    c, s = cub(interior(case["chi"])), cub(interior(case["S"]))
    gc, gs, gi = pla(case["chi"][gz]), pla(case["S"][gz]), pla(case["I"][gz])
    c500, cu, g5 = dev(case["chi500"]), dev(case["chi500_up"]), dev(case["g500"])
    Iout = torch.zeros_like(c)
    geo = dict(dir=case["dir"], ds=case["ds"], vec=case["vec"])
    ctx.sc_sweep(c, s, gc, gs, gi, c500, cu, g5, I_out=Iout, **geo)
    eb = torch.ones_like(gc)
    cb, sb = torch.zeros_like(c), torch.zeros_like(c)
    ctx.sc_adjoint(c, s, gc, gs, gi, c500, cu, g5, Iout, chibar=cb, Sbar=sb, exit_bar=eb, **geo)
    cb1, sb1 = cb.clone(), sb.clone()
    ctx.sc_adjoint(c, s, gc, gs, gi, c500, cu, g5, Iout, chibar=cb, Sbar=sb, exit_bar=eb, **geo)
    torch.cuda.synchronize()
    np.testing.assert_allclose(cb.cpu().numpy(), 2 * cb1.cpu().numpy(), rtol=1e-6)
    np.testing.assert_allclose(sb.cpu().numpy(), 2 * sb1.cpu().numpy(), rtol=1e-6)
    assert float(sb1.abs().max()) > 0
    with pytest.raises(m3drt.M3drtError, match="FP32 short_char2 only"):
        ctx.sc_adjoint(c.double(), s.double(), gc.double(), gs.double(), gi.double(), c500, cu, g5, Iout.double(),
                       chibar=cb.double(), Sbar=sb.double(), **geo)


@pytest.mark.parametrize("patch", [23, 2])
def test_formal_solution_autograd_matches_oracle(ctx, patch):
    """The torch.autograd front end (dispatch_amd.adjoint.formal_solution): the
    gradient of sum_r <W_r, exit_r> + <V, J> over four rays (two up, two down)
    against the oracle's FP64 reverse sweeps summed over the rays; with one patch
    of 5 planes, and with patches of 2, 2 and 1 planes (two checkpoint planes per
    ray, the patch-boundary cotangents handed from patch to patch)."""
    from dispatch_amd.adjoint import RaySet, formal_solution
    vecs = np.array([(0.3, -0.4, 0.866), (-0.61, 0.05, 0.79), (-0.5, 0.2, -0.84), (0.45, 0.35, -0.82)])
    vecs /= np.linalg.norm(vecs, axis=1)[:, None]
    wj = np.array([0.3, 0.2, 0.25, 0.25])
    base = make_case(nf=3, nx=8, ny=6, nz=5, seed=21, vec=vecs[0], top_ghost_zero=False)
    rng = np.random.default_rng(21)
    nz = base["nz"]
    c5 = np.concatenate([base["g500"][None], base["chi500"], (base["chi500"][-1] * 0.7)[None]]).astype(np.float64)
    cu = []
    for v in vecs:
        prev = c5[0:-2] if v[2] > 0 else c5[2:]
        cu.append((prev * np.exp(0.02 * rng.standard_normal(prev.shape))).astype(np.float32))
    chi, S = base["chi"].astype(np.float64), base["S"].astype(np.float64)
    Ig = base["I"].astype(np.float64).copy()
    Ig[-1] = S[-1] * 0.5                    # a non-zero top ghost intensity
    W = rng.uniform(-1, 1, (4,) + chi[0].shape)
    V = rng.uniform(-1, 1, chi[1:-1].shape)
    rcb, rsb = np.zeros_like(chi), np.zeros_like(chi)
    rib = np.zeros_like(chi)
    for r, v in enumerate(vecs):
        d = 1 if v[2] > 0 else -1
        g5 = (c5[0] if d > 0 else c5[-1]).astype(np.float32)
        I = O.short_char2(chi, S, Ig, base["chi500"], cu[r], g5, d, base["ds"], v, get_rates=False, fp64=True)[0]
        w = np.zeros_like(chi)
        w[1:-1] = wj[r] * V
        w[nz if d > 0 else 1] += W[r]
        Ib, cb_, sb_ = O.short_char2_adjoint(chi, S, I, base["chi500"], cu[r], g5, w, d, base["ds"], v)
        rcb += cb_
        rsb += sb_
        g = 0 if d > 0 else nz + 1
        rib[g] += Ib[g]
    tn = lambda a: dev(to_native(a))
    pn = lambda a: dev(plane_to_native(a))
    rs = RaySet(ctx, vecs, wj, base["ds"], dev(base["chi500"]), [dev(c) for c in cu],
                dev(c5[0].astype(np.float32)), dev(c5[-1].astype(np.float32)), patch=patch)
    c = tn(chi[1:-1]).requires_grad_()
    s = tn(S[1:-1]).requires_grad_()
    gu = [pn(a[0]).requires_grad_() for a in (chi, S, Ig)]
    gd = [pn(a[-1]).requires_grad_() for a in (chi, S, Ig)]
    exits, J = formal_solution(rs, c, s, gu, gd, want_J=True)
    loss = (exits * torch.from_numpy(np.moveaxis(W, 3, 1).astype(np.float32)).cuda()).sum() + (J * tn(V)).sum()
    loss.backward()
    torch.cuda.synchronize()
    back = lambda t: np.moveaxis(t.grad.double().cpu().numpy(), 1, 3)
    pback = lambda t: np.moveaxis(t.grad.double().cpu().numpy(), 0, 2)
    close(back(c), rcb[1:-1], "chi grad")
    close(back(s), rsb[1:-1], "S grad")
    close(pback(gu[0]), rcb[0], "bottom ghost chi grad")
    close(pback(gd[1]), rsb[-1], "top ghost S grad")
    close(pback(gu[2]), rib[0], "bottom ghost I grad")
    close(pback(gd[2]), rib[-1], "top ghost I grad")


def test_patch_sweeps_equal_the_column_sweep(ctx):
    """The checkpointed forward (one launch per patch of 7 planes, each patch's
    ghost plane the previous patch's last plane and exit intensity) gives the
    single-launch column sweep's exit planes and J bit for bit, up and down."""
    from dispatch_amd import synth
    from dispatch_amd.adjoint import RaySet
    box = synth.make_box(nx=24, ny=20, nz=30, nf=8, seed=31)
    chi, S = synth.build_cubes(box, torch.float32)
    gh = synth.ghost_planes(box, chi, torch.float32)
    vecs, wj = synth.gauss_rays(6)
    sel = [0, 5, 13, 20]
    rs = RaySet(ctx, vecs, wj, box.ds, box.chi500, [synth.chi500_upwind(box, v) for v in vecs], box.g500_bot,
                box.g500_top, patch=7)
    ghosts = (gh["up"], gh["down"])
    J1, J2 = torch.zeros_like(chi), torch.zeros_like(chi)
    for r in sel:
        print(f"ray {r}", flush=True)
        e1 = torch.zeros_like(gh["up"][0])
        e2 = torch.zeros_like(e1)
        rs.sweep(r, chi, S, ghosts, exit_I=e1, J=J1)
        ckpt = torch.empty((4,) + tuple(e1.shape), device="cuda")
        rs.sweep_checkpointed(r, chi, S, ghosts, ckpt, e2, J=J2)
        torch.cuda.synchronize()
        assert torch.equal(e1, e2), r
        assert float(e1.abs().max()) > 0
    assert torch.equal(J1, J2)
