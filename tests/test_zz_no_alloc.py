"""m3drt_sc_sweep_noalloc, the sweep entry point that allocates nothing (its
scratch is the library's fixed device array), against m3drt_sc_sweep on the same
inputs through the C ABI: the same kernels, so the same bits, on two calls in a
row (the second smaller), around the context's own scratch, under HIP graph
capture, and a size beyond the fixed array refused before any launch."""
import numpy as np
import pytest
import torch

from dispatch_amd import m3drt, synth

pytestmark = pytest.mark.gpu


def _problem(nx, ny, nz, nf, seed):
    box = synth.make_box(nx=nx, ny=ny, nz=nz, nf=nf, seed=seed)
    chi, S = synth.build_cubes(box, torch.float32)
    ghost = synth.ghost_planes(box, chi, torch.float32)
    rays, w = synth.gauss_rays(6)
    return box, chi, S, ghost, rays, w


def _solve(ctx, prob, rays_sel, noalloc, J=None, exits=None):
    box, chi, S, ghost, rays, w = prob
    if J is None:
        J = torch.zeros_like(chi)
    outs = []
    for r in rays_sel:
        v = rays[r]
        up = v[2] > 0
        gchi, gS, gI = ghost["up" if up else "down"]
        ex = torch.zeros_like(gchi) if exits is None else exits[r]
        ctx.sc_sweep(chi, S, gchi, gS, gI, box.chi500, synth.chi500_upwind(box, v),
                     box.g500_bot if up else box.g500_top, dir=1 if up else -1, ds=box.ds, vec=v,
                     wj=float(w[r]), J=J, exit_I=ex, noalloc=noalloc)
        outs.append(ex)
    return J, outs


RAYS = [0, 5, 13, 20]


def test_noalloc_matches_the_allocating_sweep_twice(ctx):
    # the second call is smaller: the fixed array keeps the first call's contents
    for shape, seed in (((120, 120, 46, 16), 5), ((60, 36, 20, 8), 6)):
        prob = _problem(*shape, seed)
        Ja, ea = _solve(ctx, prob, RAYS, noalloc=False)
        Jb, eb = _solve(ctx, prob, RAYS, noalloc=True)
        torch.cuda.synchronize()
        assert torch.equal(Ja, Jb)
        for a, b in zip(ea, eb):
            assert torch.equal(a, b)
        assert float(Ja.abs().max()) > 0


def test_noalloc_leaves_the_context_scratch_alone(ctx):
    prob = _problem(48, 48, 30, 8, 7)
    Ja, _ = _solve(ctx, prob, RAYS, noalloc=False)
    Jb, _ = _solve(ctx, prob, [RAYS[0]], noalloc=True)
    Jb, _ = _solve(ctx, prob, RAYS[1:], noalloc=False, J=Jb)   # the context's scratch again
    torch.cuda.synchronize()
    assert torch.equal(Ja, Jb)


def test_noalloc_under_graph_capture(ctx):
    prob = _problem(60, 60, 24, 8, 8)
    box, chi, S, ghost, rays, w = prob
    Jref, _ = _solve(ctx, prob, RAYS, noalloc=False)
    torch.cuda.synchronize()
    cus = {r: synth.chi500_upwind(box, rays[r]) for r in RAYS}
    J = torch.zeros_like(chi)
    ex = {r: torch.zeros_like(ghost["up"][0]) for r in RAYS}
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            J.zero_()
            for r in RAYS:
                v = rays[r]
                up = v[2] > 0
                gchi, gS, gI = ghost["up" if up else "down"]
                ctx.sc_sweep(chi, S, gchi, gS, gI, box.chi500, cus[r], box.g500_bot if up else box.g500_top,
                             dir=1 if up else -1, ds=box.ds, vec=v, wj=float(w[r]), J=J, exit_I=ex[r],
                             noalloc=True)
    torch.cuda.synchronize()
    for rep in range(3):
        J.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(J, Jref), rep
        # eager work between the replays, on the same fixed array
        _solve(ctx, _problem(36, 36, 12, 4, 9 + rep), RAYS[:1], noalloc=True)
        torch.cuda.synchronize()


def test_noalloc_refuses_beyond_its_fixed_size(ctx):
    # 8 bytes per cell of nx*ny*(nz+1): 120*120*301*8 = 34.7 MB > 32 MiB
    box = synth.make_box(nx=120, ny=120, nz=300, nf=1, seed=10)
    chi, S = synth.build_cubes(box, torch.float32)
    ghost = synth.ghost_planes(box, chi, torch.float32)
    v = synth.gauss_rays(6)[0][0]
    J = torch.zeros_like(chi)
    with pytest.raises(m3drt.M3drtError, match="fixed"):
        ctx.sc_sweep(chi, S, *ghost["up"], box.chi500, synth.chi500_upwind(box, v), box.g500_bot, dir=1,
                     ds=box.ds, vec=v, wj=1.0, J=J, noalloc=True)
    torch.cuda.synchronize()
    assert float(J.abs().max()) == 0.0          # nothing ran
    # the allocating entry point takes the same call
    ctx.sc_sweep(chi, S, *ghost["up"], box.chi500, synth.chi500_upwind(box, v), box.g500_bot, dir=1,
                 ds=box.ds, vec=v, wj=1.0, J=J)
    torch.cuda.synchronize()
    assert float(J.abs().max()) > 0.0
