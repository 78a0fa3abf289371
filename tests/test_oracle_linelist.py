"""Pins the oracle's line-list composition (add_linelist_opacities with
add_particle_opacity, HOP/composition_mod.f90:1663-1954, HOP/linelist_mod.f90:970-1070):

- inv_df, vbar2 and lnbase against their FP64 closed forms c/sqrt(2kT/m + v^2),
  8kT/(pi mu), ln(n scale frac inv_df) - ln U, with the reference's constants;
- logref = ln(tol chi / voigt_peak) and -huge for chi <= 0;
- one species, one isotope: the composition equals the kernel call it wraps (the
  kernel is pinned bit for bit against the reference build in test_golden_hop.py),
  and the totals are the working arrays plus the increments;
- the scattering opacity leaves in physical units, chi_scat (1 - beta).
"""
import numpy as np

from linelist_cases import make_case
from oracle import oracle as O

UU, BK, MH, CC, PI = 1.6605655e-24, 1.380662e-16, 1.6726e-24, 2.99792458e10, 3.14159265359


def test_particle_base_closed_forms():
    rng = np.random.default_rng(1)
    n = 500
    T = rng.uniform(3000, 20000, n).astype(np.float32)
    v = rng.uniform(0, 3, n).astype(np.float32)
    v[:50] = 0
    dens = 10 ** rng.uniform(-5, 14, n)
    dens[:5] = 0
    part = rng.uniform(0, 5, n)
    for awgt, frac, scale in ((55.845, 1.0, 1.0), (23.0, 0.08, 0.5)):
        idf, vb2, lnb = O.particle_base(T, v, dens, part, awgt, frac, scale)
        Td = T.astype(np.float64)
        idf_x = CC / np.sqrt(2 * BK * Td / (UU * awgt) + (v.astype(np.float64) * 1e5) ** 2)
        assert np.max(np.abs(idf / idf_x - 1)) < 3e-7
        vb2_x = 8 * BK * Td / (PI * UU) * (UU / MH + 1 / awgt)
        assert np.max(np.abs(vb2 / vb2_x - 1)) < 3e-7
        lnb_x = np.log(np.maximum(dens, np.finfo(np.float64).tiny) * scale * frac * idf_x) - part
        assert np.max(np.abs(lnb - lnb_x)) < 2e-5 * np.maximum(1, np.abs(lnb_x)).max()


def test_logref_closed_form():
    chi = np.array([1e-8, 3e-5, 0.0, -1e-9, 2.5], np.float32)
    lr = O.logref_of(chi, 1e-4)
    want = np.log(1e-4 * chi[:2].astype(np.float64) / 0.5641895835)
    assert np.max(np.abs(lr[:2] - want)) < 3e-6
    assert lr[2] == lr[3] == -np.finfo(np.float32).max


def test_single_species_is_the_kernel_call():
    c = make_case(seed=2, scat=False)
    sp = c["species"][0]
    lines, awgt, frac = sp["isotopes"][0]
    chi, eta, _, dchi, deta = O.add_linelist_opacities([sp], c["llam"], c["all_chi"], c["all_eta"], c["cells"])
    idf, vb2, lnb = O.particle_base(c["cells"]["temp"], c["cells"]["vmic"], sp["n"], sp["part"], awgt, frac)
    cl = dict(c["cells"], idf=idf, vb2=vb2)
    z = np.zeros_like(c["all_chi"])
    r = O.add_opacity_equi(lines, c["llam"], z, z, c["all_chi"], O.logref_of(c["all_chi"]), cl, lnb[None, :])
    assert np.array_equal(dchi, r["chi"]) and np.array_equal(deta, r["eta"])
    assert np.array_equal(chi, (c["all_chi"] + r["chi"]).astype(np.float32))
    assert np.array_equal(eta, (c["all_eta"] + r["eta"]).astype(np.float32))
    assert dchi.max() > 0


def test_scattering_leaves_in_physical_units():
    c = make_case(seed=3)
    args = (c["species"], c["llam"], c["all_chi"], c["all_eta"], c["cells"])
    _, _, csc, _, _ = O.add_linelist_opacities(*args, one_m_eps=c["one_m_eps"], chi_scat=c["chi_scat"], beta=None)
    _, _, csc_b, _, _ = O.add_linelist_opacities(*args, one_m_eps=c["one_m_eps"], chi_scat=c["chi_scat"],
                                                 beta=c["beta"])
    assert csc.max() > 0
    assert np.array_equal(csc_b, (csc * (np.float32(1) - c["beta"])).astype(np.float32))
