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
    with pytest.raises(m3drt.M3drtError, match="preconditioned 'largest'"):
        ctx.bb_rates(d, WMU, 1.0, g, g, precond=True, largest=True)
    with pytest.raises(m3drt.M3drtError, match="own normalisation needs wq"):
        ctx.bb_rates(dict(d, wq=None), WMU, 1.0, g, g, own_norm=True)
