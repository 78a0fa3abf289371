"""The context's scratch buffer under HIP graph capture: a captured sweep keeps the
scratch address it was recorded with, so when a later eager call needs more scratch
the captured buffer is retired (kept until the context is destroyed), not freed, and
the graph's replays stay correct.  A fresh context, a small box captured, a larger
box swept eagerly (the scratch grows), then the graph replayed with eager work
between the replays, every replay equal to the eager result."""
import pytest
import torch

from dispatch_amd import m3drt, synth

pytestmark = pytest.mark.gpu


def _problem(nx, ny, nz, nf, seed):
    box = synth.make_box(nx=nx, ny=ny, nz=nz, nf=nf, seed=seed)
    chi, S = synth.build_cubes(box, torch.float32)
    ghost = synth.ghost_planes(box, chi, torch.float32)
    rays, w = synth.gauss_rays(6)
    cus = {r: synth.chi500_upwind(box, rays[r]) for r in (0, 13)}
    return box, chi, S, ghost, rays, w, cus


def _sweeps(ctx, prob, J):
    box, chi, S, ghost, rays, w, cus = prob
    for r in (0, 13):
        v = rays[r]
        up = v[2] > 0
        gchi, gS, gI = ghost["up" if up else "down"]
        ctx.sc_sweep(chi, S, gchi, gS, gI, box.chi500, cus[r], box.g500_bot if up else box.g500_top,
                     dir=1 if up else -1, ds=box.ds, vec=v, wj=float(w[r]), J=J)


def test_captured_scratch_survives_growth():
    ctx = m3drt.Context(0)
    try:
        small = _problem(24, 20, 10, 4, 1)
        big = _problem(60, 60, 30, 8, 2)
        Jref = torch.zeros_like(small[1])
        _sweeps(ctx, small, Jref)            # sizes the scratch for the small box
        torch.cuda.synchronize()
        J = torch.zeros_like(Jref)
        s = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                J.zero_()
                _sweeps(ctx, small, J)
        torch.cuda.synchronize()
        Jbig = torch.zeros_like(big[1])
        _sweeps(ctx, big, Jbig)              # the scratch grows: the captured buffer is retired
        torch.cuda.synchronize()
        assert float(Jbig.abs().max()) > 0
        for rep in range(3):
            J.fill_(float("nan"))
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(J, Jref), rep
            _sweeps(ctx, big, Jbig)          # eager work on the grown scratch between replays
            torch.cuda.synchronize()
        del g
    finally:
        ctx.close()
