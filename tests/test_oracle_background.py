"""Pins the oracle's lte_background_emission (SRC/rt_patch_mod.f90:2127-2328): the
continuum-scattering ALI step J = (J_FS - op J_old)/(1 - op) is the exact solution
of the local problem J = op J + c in one step (J_FS = op J_old + c), keeps a fixed
point fixed, clamps op to [0, scat_op_max], falls back to J_FS where the step goes
negative (and counts it); then J holds J - B, eta = chi_cont B + wscat sigma (J - B),
dJnu = wscat (J - B); initial_Jnu = Planck gives J - B = 0."""
import numpy as np

from oracle import oracle as O


def fields(n=4000, seed=0):
    rng = np.random.default_rng(seed)
    B = rng.uniform(1e-3, 1e-1, n).astype(np.float32)
    chi = (10 ** rng.uniform(-8, -5, n)).astype(np.float32)
    sigma = (chi * rng.uniform(0.1, 5.0, n)).astype(np.float32)
    return rng, B, chi, sigma


def test_ali_step_solves_the_local_problem():
    rng, B, chi, sigma = fields()
    op = rng.uniform(0.0, 0.85, B.size)
    PsiJ = (op / sigma.astype(np.float64)).astype(np.float32)
    opf = sigma.astype(np.float64) * PsiJ.astype(np.float64)
    c = rng.uniform(1e-3, 5e-2, B.size)
    Jold = rng.uniform(1e-3, 1e-1, B.size).astype(np.float32)
    Jfs = (opf * Jold + c).astype(np.float32)
    r = O.background_emission(Jfs, B, chi, sigma, PsiJ=PsiJ, Jali_old=Jold, scat_op_max=0.9)
    want = (Jfs.astype(np.float64) - opf * Jold) / (1.0 - opf)
    np.testing.assert_allclose(r["Jali"], want.astype(np.float32), rtol=2e-7)
    # the exact local solution, up to J_FS's own FP32 rounding carried through the step
    exact = c / (1.0 - opf)
    assert np.all(np.abs(r["Jali"] - exact) <= 1.2e-7 * (Jfs / (1.0 - opf) + np.abs(exact)))
    np.testing.assert_allclose(r["J"], (r["Jali"] - B).astype(np.float32), rtol=0, atol=1e-9)
    assert r["n_neg"] == 0
    np.testing.assert_allclose(r["op_max"], opf.max(), rtol=1e-12)
    # a fixed point stays fixed
    fp = O.background_emission(Jold, B, chi, sigma, PsiJ=PsiJ, Jali_old=Jold)
    np.testing.assert_allclose(fp["Jali"], Jold, rtol=3e-7)


def test_clamp_negative_fallback_and_source():
    rng, B, chi, sigma = fields(seed=1)
    PsiJ = (rng.uniform(0.5, 3.0, B.size) / sigma).astype(np.float32)  # op beyond the clamp
    Jold = (rng.uniform(1.0, 2.0, B.size) * 1e-1).astype(np.float32)
    Jfs = (rng.uniform(1e-4, 1e-2, B.size)).astype(np.float32)          # J_FS < op J_old: negative steps
    r = O.background_emission(Jfs, B, chi, sigma, PsiJ=PsiJ, Jali_old=Jold, scat_op_max=0.7)
    assert r["op_max"] == 0.7 and r["op_raw"] > 0.7
    op = np.minimum(sigma.astype(np.float64) * PsiJ, 0.7)
    step = (Jfs - op * Jold) / (1 - op)
    neg = step < 0
    assert r["n_neg"] == int(neg.sum()) > 0
    np.testing.assert_array_equal(r["Jali"][neg], Jfs[neg])
    w = np.float32(0.5)
    s = O.background_emission(Jfs, B, chi, sigma, wscat=0.5)
    dJ = (Jfs - B).astype(np.float32)
    np.testing.assert_array_equal(s["J"], dJ)
    np.testing.assert_array_equal(s["eta"], chi * B + w * sigma * dJ)
    np.testing.assert_array_equal(s["dJnu"], w * dJ)
    p = O.background_emission(Jfs, B, chi, sigma, j_to_b=True)
    assert np.all(p["J"] == 0) and np.all(p["eta"] == chi * B)
