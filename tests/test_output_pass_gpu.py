"""The output pass at the workload's full box (120 x 120 x 230, 8 frequencies, one
inclined ray): the column solvers on the rotated cubes against the oracle run on
the same rotated inputs copied to the host (I on every plane through the exit
plane), the result files written from the GPU's exit planes against the oracle's
emergent products, and the emergent fluxes of the three solvers agree to the
accuracy of their source integrations."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_output_pass_columns_match_oracle(ctx, tmp_path):
    """... and each solver's emergent plane written through save.save_intensity
    (ie, flux of the window) reads back, as tumult's io_binary reads it, as the
    emergent products of the oracle's exit plane to 1e-5."""
    from dispatch_amd import output_pass as OP, save, synth
    from oracle import oracle as O
    from sweep_cases import max_rel
    from test_save_cpu import read_like_tumult
    nf = 8
    box = synth.make_box(nf=nf, seed=20261203)
    chi, S = synth.build_cubes(box)
    pe = OP.synthetic_pe(box)
    vecs, _ = synth.gauss_rays(6)
    vec = [float(v) for v in vecs[1]]                 # mu = 0.789, azimuth 60 degrees
    print(f"output pass ray {vec}", flush=True)
    ghosts = OP.ghost_planes(ctx, box, chi, S, pe)
    c, s, fields = OP.ray_state(ctx, box, chi, S, pe, vec, ghosts)
    ghost = ghosts["model"]
    h = lambda t: t.cpu().numpy()
    ref = lambda t: np.moveaxis(h(t), 1, 3)           # (nz, nf, ny, nx) -> (nz, ny, nx, nf)
    nz = box.nz
    work = lambda cube, plane: np.concatenate([np.moveaxis(h(plane), 0, 2)[None], ref(cube),
                                               np.zeros((1, box.ny, box.nx, nf), np.float32)], 0)
    # the ghost planes against the oracle's boundary restatements on the unrotated model
    lam_nu = synth.C_LIGHT / box.lam
    oc, oS, _ = O.lte_ghost(lam_nu, h(chi[0]).reshape(nf, -1), h(box.g500_bot), h(box.chi500[0]),
                            h(box.T_ghost_bot.float()))
    assert np.array_equal(h(ghosts["lte"][0]).reshape(nf, -1), oc)
    assert np.max(np.abs(h(ghosts["lte"][1]).reshape(nf, -1) / oS - 1)) < 1e-6
    pc = np.maximum(O.pchip_ghost(h(chi[0]), h(chi[1]), h(chi[1]), box.ds[2]), 0)
    assert np.array_equal(h(ghosts["pchip"][0]).ravel(), pc)
    for solver in OP.SOLVERS:
        print(f"  {solver}", flush=True)
        gh = OP.ghost_of(solver, ghosts)
        wchi, wS = work(c, gh[0]), work(s, gh[1])
        wI = np.zeros_like(wchi)
        wI[0] = np.moveaxis(h(gh[2]), 0, 2)
        I_out = torch.zeros_like(c)
        exit_I = torch.zeros_like(gh[0])
        kw = dict(dir=1, ds=box.ds, vec=vec, layout="native", I_out=I_out, exit_I=exit_I)
        if solver == "hermite":
            ctx.lc_sweep_hermite(c, s, *gh, fields=fields, ghost="bottom", ghost_fields=ghost, **kw)
            I, _ = O.long_char_hermite(wchi, wS, wI, ref(fields["beta"]), h(fields["pe4"]), h(fields["temp"]),
                                       h(fields["dpe_dz"]), h(fields["dlnt_dz"]), h(fields["hknu"]), 1, box.ds, vec,
                                       ghost="bottom", gpe=h(ghost["pe"]), gdpe=h(ghost["dpe_dz"]),
                                       gtemp=h(ghost["temp"]), gdlnt=h(ghost["dlnt_dz"]))
        elif solver == "hermite2":
            ctx.lc_sweep_hermite(c, s, *gh, fields=fields, ghost="bottom", ghost_fields=ghost, hybrid=True, **kw)
            I, _ = O.long_char_hermite2(wchi, wS, wI, h(fields["pe4"]), h(fields["dpe_dz"]), 1, box.ds, vec,
                                        ghost="bottom", gpe=h(ghost["pe"]), gdpe=h(ghost["dpe_dz"]))
        elif solver == "quadratic":
            ctx.sc_sweep(c, s, *gh, None, None, None, solver=solver, **kw)
            I, _ = O.long_char_quadratic(wchi, wS, wI, 1, box.ds, vec)
        else:
            ctx.sc_sweep(c, s, *gh, None, None, None, solver=solver, **kw)
            I, _, _, _ = O.long_char_linear(wchi, wS, wI, 1, box.ds, vec)
        torch.cuda.synchronize()
        Iref = I[1:-1].astype(np.float64)
        err = max_rel(ref(I_out), Iref, 1e-6 * np.abs(Iref).max())
        assert err < 1e-5, (solver, err)
        assert max_rel(np.moveaxis(h(exit_I), 0, 2), Iref[nz - 1]) < 1e-5
        # the result files of this window (one output ray of flux weight wmu)
        wmu = [0.25]
        run = str(tmp_path / solver)
        save.save_intensity(run, exit_I[None], lam_nu, wmu, key=(1, 1))
        ie = read_like_tumult(f"{run}/save", "ie")[(1, 1)]
        flux = read_like_tumult(f"{run}/save", "flux")[(1, 1)]
        ref_exit = torch.from_numpy(np.ascontiguousarray(np.moveaxis(I[nz].astype(np.float32), 2, 0)))[None]
        rie, rflux, _ = save.emergent_products(ref_exit, lam_nu, wmu)
        assert ie.shape == (nf, 1, 1, 1, 1) and flux.shape == (nf, 1, 1, 1, 1)
        assert np.max(np.abs(np.asarray(ie) / rie - 1)) < 1e-5, solver
        assert np.max(np.abs(np.asarray(flux) / rflux - 1)) < 1e-5, solver


def test_output_pass_fluxes_agree():
    from dispatch_amd import output_pass as OP
    out = OP.run(nf=16, windows=1, rays=3)
    f = out["flux_mean"]
    assert all(np.isfinite(v) and v > 0 for v in f.values())
    # the three source integrations of one atmosphere: within a few per cent
    assert all(abs(f[s] / f["quadratic"] - 1) < 0.05 for s in ("hermite", "hermite2", "long_linear"))
