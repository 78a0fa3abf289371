"""GPU parity of the line-list composition (m3drt_add_linelist_opacities:
add_linelist_opacities with add_particle_opacity, HOP/composition_mod.f90:1663-1954,
HOP/linelist_mod.f90:970-1070) against the oracle's composition of the same steps:
three species (one with two isotopes, one NLTE with departures), on an equidistant
and an uneven grid, with and without the LTE scattering split.  The Voigt sums
keep the reference's per-cell order on both sides; the per-isotope population
base takes the FP32 logarithm of inv_df, which the device forms as the FP64 one
rounded, so lnbase may differ by an ulp and the opacities by a few parts in 1e7
(tolerance 2e-6 of the largest increment)."""
import numpy as np
import pytest
import torch

from linelist_cases import make_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def d(a, dt=None):
    return torch.from_numpy(np.ascontiguousarray(a if dt is None else np.asarray(a, dt))).cuda()


def dev_species(species):
    out = []
    for sp in species:
        isos = []
        for lines, awgt, frac in sp["isotopes"]:
            dl = {k: d(v, np.float64 if k == "llam0" else (np.int32 if v.dtype.kind == "i" else np.float32))
                  for k, v in lines.items()}
            isos.append((dl, awgt, frac))
        dsp = dict(isotopes=isos, n=d(sp["n"], np.float64), part=d(sp["part"], np.float64),
                   scale=sp.get("scale", 1.0), scat_group=sp.get("scat_group", 0))
        if sp.get("dep") is not None:
            dsp["dep"] = d(sp["dep"], np.float32)
        out.append(dsp)
    return out


@pytest.mark.parametrize("equidistant,scat", [(True, True), (False, True), (True, False)])
def test_composition_matches_oracle(ctx, equidistant, scat):
    c = make_case(seed=5, equidistant=equidistant, scat=scat)
    kw = {}
    if scat:
        kw = dict(one_m_eps=c["one_m_eps"], chi_scat=c["chi_scat"], beta=c["beta"])
    chi_o, eta_o, csc_o, dchi_o, deta_o = O.add_linelist_opacities(
        c["species"], c["llam"], c["all_chi"], c["all_eta"], c["cells"], equidistant=equidistant, **kw)
    T = lambda a: d(np.ascontiguousarray(np.asarray(a, np.float32).T))          # (ncell, nf) -> (nf, ncell)
    chi, eta = T(c["all_chi"]), T(c["all_eta"])
    dchi, deta = torch.zeros_like(chi), torch.zeros_like(chi)
    dkw = {}
    if scat:
        dkw = dict(one_m_eps=d(np.ascontiguousarray(np.transpose(c["one_m_eps"], (0, 2, 1)))),
                   chi_scat=T(c["chi_scat"]), beta=T(c["beta"]))
    cells = {k: d(v, np.float32) for k, v in c["cells"].items()}
    ctx.add_linelist_opacities(dev_species(c["species"]), d(c["llam"], np.float64), chi, eta, cells,
                               equidistant=equidistant, chi_lines=dchi, eta_lines=deta, **dkw)
    torch.cuda.synchronize()
    back = lambda t: t.cpu().numpy().T
    scale = np.abs(dchi_o).max()
    assert scale > 0
    for got, want in ((back(dchi), dchi_o), (back(deta), deta_o), (back(chi), chi_o), (back(eta), eta_o)):
        assert np.max(np.abs(got - want)) <= 2e-6 * scale
    assert np.mean(back(dchi) == dchi_o) > 0.9
    if scat:
        got = back(dkw["chi_scat"])
        assert np.max(np.abs(got - csc_o)) <= 2e-6 * max(np.abs(csc_o).max(), 1e-30)
        assert np.abs(csc_o).max() > 0


@pytest.mark.parametrize("equidistant,nl,nlte", [(True, 3, False), (False, 4, False), (True, 4, True)])
def test_short_list_deposits_directly(ctx, equidistant, nl, nlte, monkeypatch, capfd):
    """A composition of one list of at most 4 lines (no other list, no scattering
    split, no copy of the increments): the summation kernel adds each point's line sum
    straight to the working opacity.  Against the oracle's composition, and bit for
    bit against the path through zeroed increments that are added at the end."""
    from linelist_cases import lines_of
    c = make_case(seed=17 + nl, equidistant=equidistant, scat=False)
    rng = np.random.default_rng(nl)
    llam = c["llam"]
    dl = (llam[-1] - llam[0]) / (len(llam) - 1)
    ncell = c["all_chi"].shape[0]
    sp = dict(isotopes=[(lines_of(rng, nl, llam, dl, 4 if nlte else 0), 40.0, 1.0)],
              n=10 ** rng.uniform(8, 12, ncell), part=rng.uniform(0.5, 4, ncell), scale=1.0, scat_group=0)
    if nlte:
        sp["dep"] = rng.uniform(0.5, 2.0, (4, ncell)).astype(np.float32)
        sp["nlev"] = 4
    species = [sp]
    print(f"direct deposit equidistant={equidistant} lines={nl} nlte={nlte}", flush=True)
    chi_o, eta_o, _, dchi_o, _ = O.add_linelist_opacities(species, llam, c["all_chi"], c["all_eta"], c["cells"],
                                                          equidistant=equidistant)
    T = lambda a: d(np.ascontiguousarray(np.asarray(a, np.float32).T))
    cells = {k: d(v, np.float32) for k, v in c["cells"].items()}
    out = []
    for env in (None, "1"):
        if env:
            monkeypatch.setenv("M3DRT_NO_DIRECT_LINES", env)
        monkeypatch.setenv("M3DRT_VERBOSE", "1")
        chi, eta = T(c["all_chi"]), T(c["all_eta"])
        ctx.add_linelist_opacities(dev_species(species), d(llam, np.float64), chi, eta, cells,
                                   equidistant=equidistant)
        torch.cuda.synchronize()
        err = capfd.readouterr().err
        assert ("deposited directly" in err) == (env is None)
        out.append((chi, eta))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    back = lambda t: t.cpu().numpy().T
    scale = np.abs(dchi_o).max()
    assert scale > 0
    assert np.max(np.abs(back(out[0][0]) - chi_o)) <= 2e-6 * scale
    assert np.max(np.abs(back(out[0][1]) - eta_o)) <= 2e-6 * scale
