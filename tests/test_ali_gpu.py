"""The NLTE ALI driver (dispatch_amd.ali) on small boxes of its workload generator:
one iteration's rate sums and departure coefficients against the oracle's
composition of the same steps (tests/ali_cases.py); convergence; the LTE limit
under strong collisional coupling."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def small(**kw):
    from dispatch_amd.ali import AliConfig
    base = dict(nx=6, ny=6, nz=48, nq=16, nbound=4, nline=6, ncont=4, lines_per_window=3, nphi=2, ncont_f=8)
    base.update(kw)
    return AliConfig(**base)


def test_ali_iteration_matches_oracle(ctx):
    from dispatch_amd.ali import NlteProblem
    from ali_cases import oracle_rates, oracle_update
    p = NlteProblem(ctx, small())
    print("ali iteration vs oracle, 6x6x48, 5 levels, 6 lines, 4 continua, 8 rays", flush=True)
    for it in range(2):
        gij, gji = p.rates()
        torch.cuda.synchronize()
        rij, rji = oracle_rates(p)
        for got, ref, what in ((gij, rij, "gamma_ij"), (gji, rji, "gamma_ji")):
            g = got.double().cpu().numpy()
            scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-300
            err = np.abs(g - ref) / scale
            assert err.max() < 2e-4, f"iteration {it} {what}: {err.max():.3e} (row {np.unravel_index(err.argmax(), err.shape)})"
        b_ref, res_ref = oracle_update(p, rij, rji)
        res = p.update(gij, gji)
        b = p.b.cpu().numpy()
        assert np.abs(b / b_ref - 1).max() < 2e-4, f"iteration {it} b: {np.abs(b / b_ref - 1).max():.3e}"
        assert abs(res["maxerr"] - res_ref["maxerr"]) <= 2e-4 * abs(res_ref["maxerr"]) + 1e-6
        p.b.copy_(torch.from_numpy(b_ref))   # both continue from the oracle's populations


def test_ali_converges_and_departs_from_lte(ctx):
    from dispatch_amd.ali import NlteProblem
    p = NlteProblem(ctx, small(max_iter=40))
    n = p.solve(log=lambda m: print(m, flush=True))
    errs = [abs(e) for e, _ in p.history]
    assert errs[-1] < 1e-4 and n <= 40, errs
    b = p.b.cpu().numpy()
    assert np.abs(b - 1).max() > 1e-2   # the lines' NLTE departures are there


def test_ali_strong_collisions_give_lte(ctx):
    """With the electron density raised 1e7-fold the lower half of the box, where
    collisions then dominate every transition, is in LTE (b within 1e-3 of 1)."""
    from dispatch_amd.ali import NlteProblem
    cfg = small(ne_scale=1e7, max_iter=40)
    p = NlteProblem(ctx, cfg)
    p.solve()
    b = p.b.cpu().numpy().reshape(p.atom.nk, cfg.nz, -1)
    assert np.abs(b[:, :cfg.nz // 2] - 1).max() < 1e-3, np.abs(b[:, :cfg.nz // 2] - 1).max()


@pytest.mark.parametrize("kw", [{}, dict(nq=50, lines_per_window=4)])
def test_window_opacity_equals_the_sequence(ctx, kw):
    """m3drt_window_opacity against the sequence it fuses (the continuum product,
    convert_to_work_units, the seeded split tracking, bb_opacity, work_to_transfer) on
    every window and direction of the small problem at non-LTE populations: chi, S and
    the six tracking arrays bit for bit (both round every product and sum apart).  With
    50 frequencies per line a window of 4 lines has 200 frequencies, which the kernel
    splits into runs of 67 per lane, their ends inside line windows."""
    from dispatch_amd.ali import NlteProblem
    from dispatch_amd import synth
    p = NlteProblem(ctx, small(**kw))
    g = torch.Generator(device="cuda").manual_seed(3)
    p.b.mul_(torch.exp(0.3 * torch.randn(p.b.shape, generator=g, dtype=torch.float64, device="cuda")))
    print(f"window opacity fused vs sequence {kw}", flush=True)
    for wi, w in enumerate(p.windows):
        l0, l1 = w.lines[0], w.lines[-1] + 1
        nf, nc = w.nf, p.ncell
        for r, v in enumerate(p.vec):
            vproj = (p.vz * (v[2] / synth.C_LIGHT)).float().contiguous()
            d = dict(w.const, vproj=vproj, b=p.b, nstar=p.nstar, pop_ratio=p.pop_ratio, inv_df=p.inv_df,
                     e_kT=p.e_kT32, adamp=p.adamp[l0:l1])
            out = []
            for fused in (False, True):
                xhat = torch.full((nf, nc), 7.0, dtype=torch.float32, device="cuda")
                ehat = torch.full_like(xhat, 7.0)
                sp = dict(split_opac1=torch.full_like(xhat, 7.0), split_opac2=torch.full_like(xhat, 7.0),
                          split_chi1=torch.full_like(xhat, 7.0), split_chi2=torch.full_like(xhat, 7.0),
                          split_id=torch.full((nf, nc), 7, dtype=torch.int32, device="cuda"),
                          split_chi_id=torch.full((nf, nc), 7, dtype=torch.int32, device="cuda"))
                if fused:
                    ctx.window_opacity(d, w.chi_c, w.Bnu, w.beta, xhat, ehat, split=sp, split_offset=0)
                else:
                    xhat.zero_()
                    torch.mul(w.chi_c, w.Bnu, out=ehat)
                    ctx.convert_to_work_units(w.chi_c, xhat, ehat, w.beta)
                    sp["split_opac1"].copy_(ehat)
                    sp["split_chi1"].copy_(xhat)
                    for k in ("split_opac2", "split_chi2", "split_id", "split_chi_id"):
                        sp[k].zero_()
                    ctx.bb_opacity(d, xhat, ehat, split=sp, split_offset=0)
                    ctx.work_to_transfer(xhat, ehat, w.beta, split_opac1=sp["split_opac1"],
                                         split_opac2=sp["split_opac2"])
                out.append(dict(sp, chi=xhat, S=ehat))
            torch.cuda.synchronize()
            for k in out[0]:
                assert torch.equal(out[0][k], out[1][k]), (wi, r, k, float((out[0][k] - out[1][k]).abs().max()))


def test_graph_replays_of_the_rate_pass_match_eager(ctx):
    """AliConfig(graph=True): the rate pass captured once and replayed, with the
    statistical-equilibrium update (eager work) between the replays, gives the eager
    pass's rate sums and departure coefficients bit for bit, iteration after
    iteration."""
    from dispatch_amd.ali import NlteProblem
    pe = NlteProblem(ctx, small())
    pg = NlteProblem(ctx, small(graph=True))
    print("ali rate pass: graph replays vs eager", flush=True)
    for it in range(4):
        ge, je = pe.rates()
        gg, jg = pg.rates()
        torch.cuda.synchronize()
        assert torch.equal(ge, gg) and torch.equal(je, jg), it
        pe.update(ge, je)
        pg.update(gg, jg)
        torch.cuda.synchronize()
        assert torch.equal(pe.b, pg.b), it
    assert pg._graph is not None
