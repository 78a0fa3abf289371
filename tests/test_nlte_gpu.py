"""GPU parity of the statistical-equilibrium update (m3drt_rate_solve_f64,
m3drt_normalize_rates) against the oracle's restatement of
setup_transition_matrix + lineq_no_pivot + update_populations.  FP64 throughout,
same operation order: departure coefficients to 1e-12 relative, the convergence
measure's cell and level exactly."""
import numpy as np
import pytest
import torch

from nlte_cases import make_se_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(a, dt):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")


@pytest.mark.parametrize("seed,ncell,kw", [(5, 1000, {}), (6, 777, dict(max_update=3.0)),
                                            (7, 300, dict(conv_pop_floor=1e-4, sor_w=0.7))])
def test_rate_solve_matches_oracle(ctx, seed, ncell, kw):
    from dispatch_amd import m3drt
    c = make_se_case(ncell=ncell, seed=seed)
    print(f"rate solve seed={seed} ncell={ncell} {kw}", flush=True)
    bref, rref = O.update_populations(c["atom"], c["e_kT"], c["T"], c["totn"], c["nstar"], c["b"], c["gij"],
                                      c["gji"], c["cup"], **kw)
    at = m3drt.AtomOnDevice(c["atom"])
    b = dev(c["b"], torch.float64)
    r = ctx.rate_solve(at, dev(c["e_kT"], torch.float64), dev(c["T"], torch.float64),
                       dev(c["totn"], torch.float64), dev(c["nstar"], torch.float64), b,
                       dev(c["gij"], torch.float32), dev(c["gji"], torch.float32),
                       dev(c["cup"], torch.float64), **kw)
    got = b.cpu().numpy()
    assert np.max(np.abs(got - bref) / np.abs(bref)) < 1e-12
    assert (r["epos"], r["elvl"]) == (rref["epos"], rref["elvl"])
    assert r["maxerr"] == pytest.approx(rref["maxerr"], rel=1e-12)
    assert r["nclamp"] == rref["nclamp"]
    assert np.allclose(r["lvlerr"], rref["lvlerr"], rtol=1e-10, atol=0)


def test_normalize_rates_matches_oracle(ctx):
    from dispatch_amd import m3drt, synth
    rng = np.random.default_rng(9)
    atom = synth.make_atom()
    ncell = 500
    T = rng.uniform(4500, 8000, ncell)
    nstar = synth.lte_populations(atom, T, 10 ** rng.uniform(12, 15, ncell), np.ones(ncell) * 1e8)
    gij = rng.uniform(0, 1, (len(atom.ti), ncell)).astype(np.float32)
    gji = rng.uniform(0, 1, (len(atom.ti), ncell)).astype(np.float32)
    wphi = rng.uniform(0.5, 2, (len(atom.ti), ncell)).astype(np.float32)
    rij, rji = O.normalize_rates(atom, nstar, gij, gji, wphi)
    dij, dji = dev(gij, torch.float32), dev(gji, torch.float32)
    ctx.normalize_rates(m3drt.AtomOnDevice(atom), dev(nstar, torch.float64), dij, dji, dev(wphi, torch.float32))
    torch.cuda.synchronize()
    assert np.array_equal(dij.cpu().numpy(), rij)
    assert np.array_equal(dji.cpu().numpy(), rji)


def test_rate_solve_lte_fixed_point(ctx):
    from dispatch_amd import m3drt
    c = make_se_case(ncell=2000, seed=10, lte=True)
    b = dev(c["b"], torch.float64)
    r = ctx.rate_solve(m3drt.AtomOnDevice(c["atom"]), dev(c["e_kT"], torch.float64), dev(c["T"], torch.float64),
                       dev(c["totn"], torch.float64), dev(c["nstar"], torch.float64), b,
                       dev(c["gij"], torch.float32), dev(c["gji"], torch.float32), dev(c["cup"], torch.float64))
    assert np.max(np.abs(b.cpu().numpy() - 1.0)) < 1e-5
    assert abs(r["maxerr"]) < 1e-5
