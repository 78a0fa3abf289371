"""GPU parity of the batched 1D synthesis (m3drt_b1d_*) against the oracle
(oracle/m3d_b1d.c, FP64 apart from the window and table-index decisions, which it
takes in FP32 as the kernel does), on a seeded batch of uneven atmospheres.

Bars: emergent intensity and flux of every (atmosphere, frequency) within 1e-5
relative of the oracle's; the plain one-launch-per-atmosphere path, every workgroup
size and every replay of the captured graph give bit-identical results (the same
per-lane arithmetic in the same order); two frequencies per lane agree with one to
1e-6 (separately compiled code)."""
import numpy as np
import pytest
import torch

from dispatch_amd import batch1d as B
from dispatch_amd import m3drt
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5
CFG = B.B1dConfig(natm=24, nd=(8, 40), nf=(30, 300), nline=(3, 30), seed=31)
TILES = [(1, 256), (1, 512), (1, 1024), (2, 256), (2, 512), (2, 1024)]


@pytest.fixture(scope="module")
def batch():
    print(f"b1d batch natm={CFG.natm}", flush=True)
    arrays, meta = B.make_batch(CFG, device="cuda")
    host = {k: v.cpu().numpy() for k, v in arrays.items()}
    ref = [O.b1d_atmosphere(host, m, B.LAM_REF) for m in range(CFG.natm)]
    return arrays, meta, ref


def split(arrays, meta, m):
    nf, nmu = int(meta["nf"][m]), int(meta["nmu"][m])
    io, fo = int(meta["i_off"][m]), int(meta["f_off"][m])
    I = arrays["I_top"][io:io + nmu * nf].view(nmu, nf).double().cpu().numpy()
    F = arrays["flux"][fo:fo + nf].double().cpu().numpy()
    return I, F


def check_oracle(arrays, meta, ref):
    worst = 0.0
    for m, (Ir, Fr) in enumerate(ref):
        I, F = split(arrays, meta, m)
        worst = max(worst, float(np.max(np.abs(I - Ir) / np.abs(Ir))), float(np.max(np.abs(F - Fr) / np.abs(Fr))))
    assert worst < TOL, worst
    return worst


def test_plan_matches_oracle(ctx, batch):
    arrays, meta, ref = batch
    arrays["I_top"].zero_()
    arrays["flux"].zero_()
    print("b1d autotuned plan", flush=True)
    plan = ctx.b1d_plan(arrays, B.LAM_REF)
    tile, times = plan.info()
    assert tile in TILES and len(times) == 6 and all(t > 0 for t in times)
    plan.run()
    torch.cuda.synchronize()
    check_oracle(arrays, meta, ref)
    plan.close()


def test_plain_and_every_tile_are_identical(ctx, batch):
    arrays, meta, ref = batch
    print("b1d plain path, tile (1, 256)", flush=True)
    ctx.b1d_synth_plain(arrays, B.LAM_REF)
    torch.cuda.synchronize()
    check_oracle(arrays, meta, ref)
    I0, F0 = arrays["I_top"].clone(), arrays["flux"].clone()
    for t in TILES:
        print(f"b1d tile {t}", flush=True)
        for run in ("plain", "plan"):
            arrays["I_top"].fill_(-1.0)
            arrays["flux"].fill_(-1.0)
            if run == "plain":
                ctx.b1d_synth_plain(arrays, B.LAM_REF, tile=t)
            else:
                p = ctx.b1d_plan(arrays, B.LAM_REF, tile=t)
                assert p.info()[0] == t
                p.run()
            torch.cuda.synchronize()
            if t[0] == 1:
                assert torch.equal(arrays["I_top"], I0) and torch.equal(arrays["flux"], F0), (t, run)
            else:
                torch.testing.assert_close(arrays["I_top"], I0, rtol=1e-6, atol=0)
                torch.testing.assert_close(arrays["flux"], F0, rtol=1e-6, atol=0)
                if t[1] == 256 and run == "plain":
                    I2, F2 = arrays["I_top"].clone(), arrays["flux"].clone()
                else:
                    assert torch.equal(arrays["I_top"], I2) and torch.equal(arrays["flux"], F2), (t, run)
            if run == "plan":
                p.close()


def test_graph_replays_follow_the_inputs(ctx, batch):
    """Replays of one plan with eager work between them: each replay re-forms the
    profile records from line_rec and gives what a fresh plain run gives."""
    arrays, meta, ref = batch
    plan = ctx.b1d_plan(arrays, B.LAM_REF, tile=(1, 256))
    rec0 = arrays["line_rec"].clone()
    outs = []
    for k, scale in enumerate((1.0, 3.0, 1.0, 0.5)):
        print(f"b1d replay {k} strength x{scale}", flush=True)
        arrays["line_rec"].copy_(rec0)
        arrays["line_rec"].view(-1, 4)[:, 1] *= scale
        arrays["flux"].fill_(-1.0)
        plan.run()
        torch.cuda.synchronize()
        got = arrays["flux"].clone()
        ctx.b1d_synth_plain(arrays, B.LAM_REF, tile=(1, 256))
        torch.cuda.synchronize()
        assert torch.equal(got, arrays["flux"]), k
        outs.append(got)
    assert torch.equal(outs[0], outs[2]) and not torch.equal(outs[0], outs[1])
    arrays["line_rec"].copy_(rec0)
    plan.close()


def test_rejects_bad_tiles_and_counts(ctx, batch):
    arrays, meta, ref = batch
    with pytest.raises(m3drt.M3drtError):
        ctx.b1d_synth_plain(arrays, B.LAM_REF, tile=(3, 256))
    with pytest.raises(m3drt.M3drtError):
        ctx.b1d_plan(arrays, B.LAM_REF, tile=(1, 384))
    bad = dict(arrays, nmu=torch.full_like(arrays["nmu"], 9))
    with pytest.raises(ValueError):
        ctx.b1d_synth_plain(bad, B.LAM_REF)


def test_single_atmosphere_synthesis_matches_oracle(ctx):
    """The 72-point, 4000-wavelength, 60-line atmosphere with 5 Gauss-Radau mu through
    the plan, against the oracle in FP64."""
    print("b1d config0", flush=True)
    arrays, meta = B.make_batch(B.CONFIG0, device="cuda")
    plan = ctx.b1d_plan(arrays, B.LAM_REF, tile=(1, 256))
    plan.run()
    torch.cuda.synchronize()
    host = {k: v.cpu().numpy() for k, v in arrays.items()}
    check_oracle(arrays, meta, [O.b1d_atmosphere(host, 0, B.LAM_REF)])
    plan.close()


def test_full_workload_sample_matches_oracle(ctx):
    """The 8192-atmosphere workload itself (depths 48-256, 200-6000 frequencies, 3-7
    mu, 20-200 lines) through the autotuned persistent plan, and the oracle on its
    atmospheres with the most depth points, frequencies, mu points and
    lines, the most (depth, frequency, mu) updates, and 24 more drawn at random."""
    arrays, meta = B.make_batch(B.B1dConfig(), device="cuda")
    print("b1d full workload, 8192 atmospheres", flush=True)
    plan = ctx.b1d_plan(arrays, B.LAM_REF)
    plan.run()
    torch.cuda.synchronize()
    plan.close()
    upd = meta["nd"].astype(np.int64) * meta["nf"] * meta["nmu"]
    pick = {int(np.argmax(meta[k])) for k in ("nd", "nf", "nmu", "nline")} | {int(np.argmax(upd))}
    pick |= set(int(m) for m in np.random.default_rng(3).choice(len(meta["nd"]), 24, replace=False))
    host = {k: v.cpu().numpy() for k, v in arrays.items()}
    worst = 0.0
    for m in sorted(pick):
        print(f"  atmosphere {m}: nd {meta['nd'][m]} nf {meta['nf'][m]} nmu {meta['nmu'][m]} "
              f"lines {meta['nline'][m]}", flush=True)
        Ir, Fr = O.b1d_atmosphere(host, m, B.LAM_REF)
        I, F = split(arrays, meta, m)
        worst = max(worst, float(np.max(np.abs(I - Ir) / np.abs(Ir))), float(np.max(np.abs(F - Fr) / np.abs(Fr))))
    assert worst < TOL, worst
    assert max(meta["nd"][m] for m in pick) > 200 and max(meta["nf"][m] for m in pick) > 5000
