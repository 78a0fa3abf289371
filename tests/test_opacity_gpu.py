"""GPU parity of the line-opacity summation (m3drt_line_opacity_equi) against
add_opacity_equi: the reference's own compiled outputs (tests/golden) and the
oracle's restatement on a larger seeded case.  FP32; the per-point reciprocal
and device exp/log differ from the reference's by rounding only.  Tolerance:
2e-6 relative on the deposited opacity (the reference's golden opacity tests
use 1e-5), exact on the set of lines used.
"""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LINE_KEYS_F = ("lower_ev", "gpref", "grad", "gw1", "gw2", "gstk")
LINE_KEYS_I = ("unsold", "lev_i", "lev_j")
CELL_KEYS = ("vproj", "idf", "vb2", "ekt", "ne", "tothi", "spe")


def dev(a, dt=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t.to(dt) if dt is not None else t).cuda()


def run(ctx, g, layout, flags, general=False):
    lines = {"llam0": dev(g["in_llam0"], torch.float64)}
    lines.update({k: dev(g["in_" + k], torch.float32) for k in LINE_KEYS_F})
    lines.update({k: dev(g["in_" + k], torch.int32) for k in LINE_KEYS_I})
    if flags & 2:
        lines["ipart"] = dev(g["in_ipart"], torch.int32)
    if flags & 8:
        lines["reach_in"] = dev(g["in_reach_in"], torch.float32)
        lines["active"] = dev(g["in_active"], torch.int32)
    sp = (lambda a: dev(a.T)) if layout == "native" else (lambda a: dev(a))
    chi, eta = sp(g["in_chi"]), sp(g["in_eta"])
    ome = sp(g["in_one_m_eps"]) if flags & 4 else None
    csc = sp(g["in_chi_scat"]) if flags & 4 else None
    nl = len(g["in_llam0"])
    used = torch.zeros(nl, dtype=torch.int32, device="cuda")
    strength = torch.zeros(nl, dtype=torch.float32, device="cuda")
    reach = torch.zeros(nl, dtype=torch.float32, device="cuda")
    ctx.line_opacity_equi(lines, dev(g["in_llam"], torch.float64), chi, eta, sp(g["in_cont"]),
                          sp(g["in_logref"]), {k: dev(g["in_" + k]) for k in CELL_KEYS},
                          dev(g["in_lnbase"]), dep=dev(g["in_dep"]), nlte=bool(flags & 1),
                          one_m_eps=ome, chi_scat=csc, layout=layout, used=used, strength=strength,
                          reach=reach, general=general)
    torch.cuda.synchronize()
    back = (lambda t: t.cpu().numpy().T) if layout == "native" else (lambda t: t.cpu().numpy())
    out = dict(chi=back(chi), eta=back(eta), used=used.cpu().numpy(), strength=strength.cpu().numpy(),
               reach=reach.cpu().numpy())
    if csc is not None:
        out["chi_scat"] = back(csc)
    return out


@pytest.mark.parametrize("name", ["hop_opac_equi_lte.npz", "hop_opac_equi_nlte.npz",
                                  "hop_opac_equi_scat.npz", "hop_opac_gen_lte.npz", "hop_opac_gen_nlte.npz",
                                  "hop_opac_gen_scat.npz", "hop_opac_equi_reach.npz",
                                  "hop_opac_equi_reach_scat.npz", "hop_opac_gen_reach.npz",
                                  "hop_opac_gen_reach_nlte.npz"])
@pytest.mark.parametrize("layout", ["native", "reference"])
def test_opacity_matches_reference_build(ctx, name, layout):
    """m3drt_line_opacity_equi on the equidistant fixtures, m3drt_line_opacity
    (add_opacity) on the hop_opac_gen ones, whose grid steps vary by a factor 5; the
    _reach fixtures are lists with reaches and active flags (select_lines, the lines
    beyond the grid with their capped windows)."""
    g = np.load(os.path.join(GOLD, name))
    flags = int(g["in_flags"])
    print(f"opacity {name} {layout}", flush=True)
    r = run(ctx, g, layout, flags, general="_gen_" in name)
    assert np.array_equal(r["used"], g["used"])
    for k in ("chi", "eta", "strength", "reach") + (("chi_scat",) if flags & 4 else ()):
        err = np.max(np.abs(r[k] - g[k]) / np.maximum(np.abs(g[k]), 1e-30))
        assert err < 2e-6, f"{k}: {err:.2e}"


def test_opacity_matches_oracle_many_cells(ctx):
    rng = np.random.default_rng(31)
    g = dict(np.load(os.path.join(GOLD, "hop_opac_equi_nlte.npz")))
    # tile the fixture's cells 64 times with jittered thermodynamics
    rep = 64
    for k in CELL_KEYS:
        g["in_" + k] = (np.tile(g["in_" + k], rep) * rng.uniform(0.97, 1.03, g["in_" + k].size * rep)
                        ).astype(np.float32)
    for k in ("in_chi", "in_eta", "in_cont", "in_logref"):
        g[k] = np.tile(g[k], (rep, 1))
    g["in_lnbase"] = (np.tile(g["in_lnbase"], (1, rep)) + rng.uniform(-1, 1, (2, g["in_lnbase"].shape[1] * rep))
                      ).astype(np.float32)
    g["in_dep"] = np.tile(g["in_dep"], (1, rep))
    lines = {k: g["in_" + k] for k in ("llam0",) + LINE_KEYS_F + LINE_KEYS_I}
    lines["ipart"] = g["in_ipart"]
    ref = O.add_opacity_equi(lines, g["in_llam"], g["in_chi"], g["in_eta"], g["in_cont"], g["in_logref"],
                             {k: g["in_" + k] for k in CELL_KEYS}, g["in_lnbase"], dep=g["in_dep"],
                             do_nlte=True)
    r = run(ctx, g, "native", 3)
    assert np.array_equal(r["used"], ref["used"])
    for k in ("chi", "eta"):
        err = np.max(np.abs(r[k] - ref[k]) / np.maximum(np.abs(ref[k]), 1e-30))
        assert err < 2e-6, f"{k}: {err:.2e}"


def test_missing_selection_scratch_is_an_error(ctx):
    from dispatch_amd import m3drt
    g = np.load(os.path.join(GOLD, "hop_opac_equi_reach.npz"))
    nl = len(g["in_llam0"])
    lines = {"llam0": dev(g["in_llam0"], torch.float64)}
    lines.update({k: dev(g["in_" + k], torch.float32) for k in LINE_KEYS_F})
    lines.update({k: dev(g["in_" + k], torch.int32) for k in LINE_KEYS_I})
    lines["active"] = dev(g["in_active"], torch.int32)
    L = m3drt.lib()
    ll = m3drt.Linelist(nl, *[m3drt.C.c_void_p(lines[k].data_ptr()) for k in (
        "llam0", "lower_ev", "gpref", "grad", "gw1", "gw2", "gstk", "unsold", "lev_i", "lev_j")],
        None, None, None, None, None, m3drt.C.c_void_p(lines["active"].data_ptr()), None)
    o = m3drt.OpacityArgs()
    o.nf, o.ncell, o.npart, o.sf, o.sc = 4, 1, 1, 1, 4
    assert L.m3drt_line_opacity_equi(ctx._h, m3drt.C.byref(ll), m3drt.C.byref(o)) != 0
    assert "sel_work" in m3drt.last_error()
