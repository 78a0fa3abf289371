"""The oracle against golden vectors from the reference's own compiled code.

tests/golden/*.npz were written by tests/golden/make_golden.py, which runs
oracle/ref_harness/hop_golden.f90 linked against HOP objects compiled from the
reference's sources (amdflang -O2, no fast-math; provenance inside each file).
The restatement follows the same FP32 operation order, so most quantities agree
bit for bit; where a different compiler's choices can move the last bit the
tolerance is stated.  Independently of the reference, qlg_ref_f64 is checked
against scipy's Faddeeva function and QLG-bot against the accuracy the reference
states for it (HOP/qlg_mod.f90:15-18: 5.7e-3).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def test_exp_neg_matches_reference_build():
    g = load("hop_expneg.npz")
    E, u0 = O.exp_neg(g["in_t"])
    assert np.array_equal(E, g["E"])
    assert np.array_equal(u0, g["u0"])


def test_qlg_bot_matches_reference_build():
    g = load("hop_qlg.npz")
    K = O.qlg_bot(g["in_a"], g["in_v"])
    assert np.array_equal(K, g["K"])


def test_qlg_ref_f64_matches_reference_build():
    g = load("hop_qlg.npz")
    K = O.qlg_ref_f64(g["in_ad"], g["in_vd"])
    # the Weideman coefficients are recomputed by the published DFT recipe, so the
    # smallest ones differ from the reference's tabulated values in their last digits
    assert np.allclose(K, g["Kref"], rtol=1e-9, atol=1e-15)


def test_qlg_ref_f64_is_faddeeva():
    special = pytest.importorskip("scipy.special")
    rng = np.random.default_rng(7)
    a = 10 ** rng.uniform(-4, 2, 3000)
    v = rng.uniform(-40, 40, 3000)
    K = O.qlg_ref_f64(a, v)
    exact = special.wofz(v + 1j * a).real
    # Humlicek's region-I asymptote beyond |v| + a = 15 is good to ~1e-4 there
    assert np.max(np.abs(K - exact) / exact) < 1.2e-4
    inner = np.abs(v) + a < 15
    # Weideman N = 32: ~1e-14 absolute
    assert np.max(np.abs(K[inner] - exact[inner])) < 1e-13
    assert np.max(np.abs(K[inner] - exact[inner]) / exact[inner]) < 1e-7


def test_qlg_bot_accuracy_against_faddeeva():
    special = pytest.importorskip("scipy.special")
    rng = np.random.default_rng(8)
    a = (10 ** rng.uniform(-4, 2, 20000)).astype(np.float32)
    v = rng.uniform(0, 12, 20000).astype(np.float32)
    K = O.qlg_bot(a, v).astype(np.float64)
    exact = special.wofz(v.astype(np.float64) + 1j * a.astype(np.float64)).real
    rel = np.abs(K - exact) / exact
    # 5.7e-3 is the figure over the paper's validation domain; this wider random
    # sample (a to 100, v to 12) peaks at 6.2e-3.  A wrong fitted constant or a
    # table built on the wrong grid is off by more than 1e-2.
    assert rel.max() < 6.5e-3
    assert np.median(rel) < 3e-3


def test_qlg_line_sweep_matches_reference_build():
    g = load("hop_qlg.npz")
    llam = g["in_llam"]
    for c in range(len(g["line_fa"])):
        D, gL, S, ref, lnq, cap = (np.float32(x) for x in g["in_case_f32"][c])
        llc, pos, idl = g["in_case_f64"][c]
        fa, fb, hw, K = O.qlg_line_sweep(D, gL, S, ref, lnq, llam, llc, pos, idl, cap)
        assert (fa, fb) == (g["line_fa"][c], g["line_fb"][c])
        assert np.float32(hw) == g["line_hw"][c]
        n = fb - fa + 1
        if n > 0:
            assert np.array_equal(K[:n], g["line_K"][c][:n])


def test_lineq_no_pivot_reference_solution():
    g = load("hop_lineq.npz")
    for A, b, x in zip(g["in_A"], g["in_b"], g["x"]):
        assert np.allclose(A @ x, b, rtol=0, atol=1e-10 * np.abs(b).max())
        assert np.allclose(np.linalg.solve(A, b), x, rtol=1e-9, atol=1e-12)


OPAC_GOLDENS = ["hop_opac_equi_lte.npz", "hop_opac_equi_nlte.npz", "hop_opac_equi_scat.npz",
                "hop_opac_gen_lte.npz", "hop_opac_gen_nlte.npz", "hop_opac_gen_scat.npz",
                "hop_opac_equi_reach.npz", "hop_opac_equi_reach_scat.npz", "hop_opac_gen_reach.npz",
                "hop_opac_gen_reach_nlte.npz"]


@pytest.mark.parametrize("name", OPAC_GOLDENS)
def test_add_opacity_equi_matches_reference_build(name):
    """add_opacity_equi, and add_opacity on grids whose steps vary by a factor 5
    (the hop_opac_gen fixtures), against the reference's compiled outputs; the
    _reach fixtures hold lists with reaches and active flags (select_lines'
    narrowing and per-line test, the lines centred beyond the grid with the grid
    end's reference and their windows capped by the reach)."""
    g = load(name)
    fn = O.add_opacity if "_gen_" in name else O.add_opacity_equi
    flags = int(g["in_flags"])
    lines = {k: g["in_" + k] for k in ("llam0", "lower_ev", "gpref", "grad", "gw1", "gw2", "gstk",
                                        "unsold", "lev_i", "lev_j")}
    lines["ipart"] = g["in_ipart"] if flags & 2 else None
    if flags & 8:
        lines["reach_in"] = g["in_reach_in"]
        lines["active"] = g["in_active"]
        # the case exercises what the reaches change: inactive lines, lines outside
        # the grid that deposit, lines the selection leaves out
        llam = g["in_llam"]
        out = (g["in_llam0"] < llam[0]) | (g["in_llam0"] > llam[-1])
        assert np.any(g["in_active"] == 0) and not np.any(g["used"][g["in_active"] == 0])
        assert np.any(g["used"][out] == 1)
        assert np.any((g["reach"] == 0) & (g["in_active"] == 1) & out)
    cells = {k: g["in_" + k] for k in ("vproj", "idf", "vb2", "ekt", "ne", "tothi", "spe")}
    r = fn(lines, g["in_llam"], g["in_chi"], g["in_eta"], g["in_cont"], g["in_logref"],
           cells, g["in_lnbase"], dep=g["in_dep"], do_nlte=bool(flags & 1),
           one_m_eps=g["in_one_m_eps"] if flags & 4 else None,
           chi_scat=g["in_chi_scat"] if flags & 4 else None)
    assert np.array_equal(r["used"], g["used"])
    assert r["used"].sum() > 5                                   # the case exercises the kernel
    assert np.allclose(r["chi"], g["chi"], rtol=2e-6, atol=0)
    assert np.allclose(r["eta"], g["eta"], rtol=2e-6, atol=0)
    assert np.allclose(r["strength"], g["strength"], rtol=2e-6, atol=0)
    assert np.allclose(r["reach"], g["reach"], rtol=2e-6, atol=0)
    if flags & 4:
        assert np.allclose(r["chi_scat"], g["chi_scat"], rtol=2e-6, atol=0)
    # the line opacity deposited is the sum of QLG profiles: positive where touched
    assert np.all(r["chi"] >= g["in_chi"])
