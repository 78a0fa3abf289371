"""GPU parity of the line-rate integrals (m3drt_bb_rates) against the oracle's
restatement of bound_bound_rates with the same Kahan-compensated FP32 sums.

The kernel and the oracle evaluate the same profile arithmetic (QLG-bot tables,
contraction off) and the same compensated sums in the same frequency order, so
they agree to a few FP32 roundings; the bar is 2e-6 relative with a floor of 1e-6
of the largest value (dev cancels to near zero where I is close to Bnu).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rates_cases import make_rates_case
from dispatch_amd import m3drt

pytestmark = pytest.mark.gpu
WMU = 1.0 / 24


def to_dev(d, layout):
    out = {}
    for k, v in d.items():
        a = np.ascontiguousarray(v)
        if layout == "reference" and k in ("Inu", "Psi", "Bnu", "beta"):
            a = np.ascontiguousarray(a.T)
        out[k] = torch.from_numpy(a).to(m3drt.BB_DTYPES[k]).cuda()
    return out


def check(got, ref, what, rtol=2e-6):
    ref = np.asarray(ref, np.float64)
    got = np.asarray(got, np.float64)
    floor = 1e-6 * np.max(np.abs(ref))
    err = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), floor))
    assert err < rtol, f"{what}: max rel err {err:.3e}"


KW = [dict(), dict(precond=True), dict(own_norm=True), dict(largest=True), dict(precond=True, own_norm=True)]


@pytest.mark.parametrize("layout", ["native", "reference"])
@pytest.mark.parametrize("kw", KW)
def test_bb_rates_match_oracle(ctx, kw, layout):
    d = make_rates_case(ncell=700, nf=240, nline=6, seed=11)
    print(f"bb_rates {kw} {layout}", flush=True)
    rng = np.random.default_rng(1)
    g0 = rng.uniform(-1.0, 1.0, d["adamp"].shape).astype(np.float32) * 1e20
    ref = O.bb_rates(d, WMU, 0.9, g0, g0, kahan=True, **kw)
    dd = to_dev(d, layout)
    gij = torch.from_numpy(g0.copy()).cuda()
    gji = torch.from_numpy(g0.copy()).cuda()
    ctx.bb_rates(dd, WMU, 0.9, gij, gji, layout=layout, **kw)
    torch.cuda.synchronize()
    check(gij.cpu().numpy(), ref[0], "gamma_ij (th)")
    check(gji.cpu().numpy(), ref[1], "gamma_ji (dev)")


def test_bb_rates_two_directions_and_empty(ctx):
    """Two directions accumulate; a call with no lines is a no-op."""
    d = make_rates_case(ncell=130, nf=120, nline=4, seed=12)
    z = np.zeros(d["adamp"].shape, np.float32)
    r1 = O.bb_rates(d, WMU, 1.0, z, z, kahan=True, precond=True)
    d2 = dict(d, Inu=(d["Inu"] * 1.1).astype(np.float32), vproj=(-d["vproj"]).astype(np.float32))
    r2 = O.bb_rates(d2, 2 * WMU, 1.0, r1[0], r1[1], kahan=True, precond=True)
    gij = torch.zeros(d["adamp"].shape, dtype=torch.float32, device="cuda")
    gji = torch.zeros_like(gij)
    ctx.bb_rates(to_dev(d, "native"), WMU, 1.0, gij, gji, precond=True)
    ctx.bb_rates(to_dev(d2, "native"), 2 * WMU, 1.0, gij, gji, precond=True)
    torch.cuda.synchronize()
    check(gij.cpu().numpy(), r2[0], "two directions th")
    check(gji.cpu().numpy(), r2[1], "two directions dev")
    e = {k: (v[:0] if k in ("iblue", "ired", "lev_i", "lev_j", "Aij", "Bij", "lower_off") else v)
         for k, v in to_dev(d, "native").items()}
    e["q"] = e["q"][:0]
    e["wnu"], e["wq"], e["adamp"] = e["wnu"][:0], e["wq"][:0], e["adamp"][:0]
    empty = torch.zeros((0, 130), dtype=torch.float32, device="cuda")
    ctx.bb_rates(e, WMU, 1.0, empty, empty)


def test_bb_rates_rejects_bad_arguments(ctx):
    d = to_dev(make_rates_case(ncell=64, nf=60, nline=3, seed=13), "native")
    g = torch.zeros((3, 64), dtype=torch.float32, device="cuda")
    with pytest.raises(m3drt.M3drtError, match="'largest' with preconditioning needs"):
        ctx.bb_rates(d, WMU, 1.0, g, g, precond=True, largest=True)
    with pytest.raises(m3drt.M3drtError, match="own normalisation needs wq"):
        ctx.bb_rates(dict(d, wq=None), WMU, 1.0, g, g, own_norm=True)


def dev_dict(d, dtypes=None):
    dtypes = dtypes or m3drt.BB_DTYPES
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(dtypes[k]).cuda() for k, v in d.items()}


def zeros_split(nf, ncell):
    z = np.zeros((nf, ncell), np.float32)
    zi = np.zeros((nf, ncell), np.int32)
    return dict(split_opac1=z, split_opac2=z, split_chi1=z, split_chi2=z, split_id=zi, split_chi_id=zi)


@pytest.mark.parametrize("lte", [False, True])
def test_bb_opacity_matches_oracle(ctx, lte):
    """add_bound_bound_opacities with the split tracking and wphi: the profile
    arithmetic and the order of the additions are the oracle's, so xhat, ehat, the
    tracked terms and their indices agree to the last bit except where the FP64
    wphi sum rounds differently (2e-7)."""
    from rates_cases import make_rates_case
    d = make_rates_case(ncell=300, nf=200, nline=6, seed=31)
    nline, nf = d["q"].shape
    ncell = d["b"].shape[1]
    rng = np.random.default_rng(3)
    x0 = rng.uniform(0, 1e-3, (nf, ncell)).astype(np.float32)
    e0 = rng.uniform(0, 1e-3, (nf, ncell)).astype(np.float32)
    sp0 = zeros_split(nf, ncell)
    wl = (np.arange(nline) % 2).astype(np.int32)
    w0 = np.full((nline, ncell), 0.25, np.float32)
    ref = O.bb_opacity(d, x0, e0, wmu=0.3, lte=lte, split=sp0, split_offset=5, wphi_line=wl, wphi=w0)
    g = dev_dict(dict(xhat=x0, ehat=e0, wphi=w0, wphi_line=wl, **sp0))
    ctx.bb_opacity(dev_dict(d), g["xhat"], g["ehat"], wmu=0.3, lte=lte,
                   split={k: g[k] for k in sp0}, split_offset=5, wphi_line=g["wphi_line"], wphi=g["wphi"])
    torch.cuda.synchronize()
    for k in ("xhat", "ehat", "split_opac1", "split_opac2", "split_chi1", "split_chi2", "split_id", "split_chi_id"):
        np.testing.assert_array_equal(g[k].cpu().numpy(), ref[k], err_msg=k)
    np.testing.assert_allclose(g["wphi"].cpu().numpy(), ref["wphi"], rtol=2e-7)


@pytest.mark.parametrize("seed,scat,flags", [(0, False, False), (1, True, True)])
def test_bb_rates_split_preconditioned_matches_oracle(ctx, seed, scat, flags):
    from rates_cases import make_equal_pair_case, split_solver_outputs
    d = make_equal_pair_case(seed=seed, ncell=400)
    nline, nf = d["q"].shape
    ncell = d["b"].shape[1]
    z = np.zeros((nf, ncell), np.float32)
    opac = O.bb_opacity(d, z, z, split=zeros_split(nf, ncell), split_offset=3)
    chi = opac["xhat"] - opac["ehat"] * d["beta"]
    rng = np.random.default_rng(seed + 7)
    Psi = np.where(chi > 0, rng.uniform(0.05, 0.8, chi.shape) / np.where(chi > 0, chi, 1.0), 0.0).astype(np.float32)
    Psi[:, ::17] = 0.0
    d["Psi"] = Psi
    sp = split_solver_outputs(d, opac, Psi)
    if scat:
        sp["chi_scat"] = rng.uniform(0, 1e-3, (nf, ncell)).astype(np.float32)
        sp["dJnu"] = rng.uniform(-0.1, 0.1, (nf, ncell)).astype(np.float32)
    if flags:
        sp["ntrans"] = (np.arange(nline, dtype=np.int32) * 13 + 2)
        npw = (int(sp["ntrans"].max()) - 1) // 32 + 1
        sp["psi_flags_in"] = rng.integers(-2**31, 2**31 - 1, (npw, ncell), dtype=np.int64).astype(np.int32)
    g0 = rng.uniform(-1, 1, (nline, ncell)).astype(np.float32) * 1e10
    ref = O.bb_rates(d, WMU, 0.8, g0, g0, precond=True, largest=True, own_norm=True, kahan=True, split=sp,
                     split_offset=3)
    gsp = dev_dict(sp)
    if flags:
        gsp["psi_flags_out"] = torch.zeros_like(gsp["psi_flags_in"])
    gij = torch.from_numpy(g0.copy()).cuda()
    gji = torch.from_numpy(g0.copy()).cuda()
    ctx.bb_rates(dev_dict(d), WMU, 0.8, gij, gji, precond=True, largest=True, own_norm=True, split=gsp,
                 split_offset=3)
    torch.cuda.synchronize()
    check(gij.cpu().numpy(), ref[0], "split gamma_ij")
    check(gji.cpu().numpy(), ref[1], "split gamma_ji")
    if flags:
        np.testing.assert_array_equal(gsp["psi_flags_out"].cpu().numpy(), ref[2])


def test_bf_rates_match_oracle(ctx):
    from test_oracle_ali import bf_case
    d, nline = bf_case(33, ncell=500)
    ncont, nf = d["alpha"].shape
    ncell = d["nstar"].shape[1]
    rng = np.random.default_rng(4)
    d["Jnu"] = (d["Jnu"] * rng.uniform(0.5, 1.5, d["Jnu"].shape)).astype(np.float32)
    d["PsiJ"] = rng.uniform(0, 1e-3, (nf, ncell)).astype(np.float32)
    d["active"] = np.array([1, 0, 1], np.int32)
    g0 = np.full((nline + ncont, ncell), 0.5, np.float32)
    for precond in (False, True):
        ref = O.bf_rates(d, nline, g0, g0, precond=precond, cut_x=15.0)
        gij = torch.from_numpy(g0.copy()).cuda()
        gji = torch.from_numpy(g0.copy()).cuda()
        ctx.bf_rates(dev_dict(d, m3drt.BF_DTYPES), nline, gij, gji, precond=precond, cut_x=15.0)
        torch.cuda.synchronize()
        # the same FP32 sums in the same order: bit for bit
        np.testing.assert_array_equal(gij.cpu().numpy(), ref[0])
        np.testing.assert_array_equal(gji.cpu().numpy(), ref[1])


@pytest.mark.parametrize("kw", [dict(), dict(precond=True, own_norm=True)])
def test_bb_rates_line_runs_equal_single_lines(ctx, kw):
    """Over 300 000 cells the kernel gives each lane a run of lines (8 lines in runs of
    2); each (line, cell) sum is its own, so the call equals one call per line bit for
    bit."""
    d = make_rates_case(ncell=300_000, nf=64, nline=8, seed=5)
    print(f"bb_rates line runs {kw}", flush=True)
    per_line = ("iblue", "ired", "lev_i", "lev_j", "Aij", "Bij", "lower_off", "q", "wnu", "wq", "adamp")
    dd = to_dev(d, "native")
    g0 = torch.zeros_like(dd["adamp"])
    full = (g0.clone(), g0.clone())
    ctx.bb_rates(dd, WMU, 0.9, full[0], full[1], **kw)
    each = (g0.clone(), g0.clone())
    for l in range(8):
        dl = {k: (v[l:l + 1] if k in per_line else v) for k, v in dd.items()}
        ctx.bb_rates(dl, WMU, 0.9, each[0][l:l + 1], each[1][l:l + 1], **kw)
    torch.cuda.synchronize()
    assert float(full[0].abs().max()) > 0
    assert torch.equal(full[0], each[0]) and torch.equal(full[1], each[1])
