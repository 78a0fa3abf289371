"""GPU parity of the working-unit conversions (m3drt_convert_to_work_units,
m3drt_work_to_transfer) against the oracle: the same FP32 operations in the same
order per point, so bit for bit, and the same count of floored points."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from test_oracle_units import fields

pytestmark = pytest.mark.gpu


def d(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.mark.parametrize("mode", ["plain", "split", "scat", "suppress"])
def test_units_match_oracle(ctx, mode):
    rng, beta, cont, eta = fields(n=300001, seed=5)
    xhat, ehat = O.convert_to_work_units(cont, eta, beta)
    gx, ge = torch.zeros_like(d(cont)), d(eta)
    ctx.convert_to_work_units(d(cont), gx, ge, d(beta))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(gx.cpu().numpy(), xhat)
    np.testing.assert_array_equal(ge.cpu().numpy(), ehat)
    # push some points through the maser floor
    ehat = (ehat * np.where(rng.uniform(size=ehat.size) < 0.2, rng.uniform(0.9, 1.2, ehat.size) * xhat / (ehat * beta),
                            1.0)).astype(np.float32)
    kw = {}
    if mode == "split":
        so1 = (ehat * rng.uniform(0.2, 0.8, ehat.size)).astype(np.float32)
        kw = dict(split_opac1=so1, split_opac2=(ehat - so1).astype(np.float32))
    if mode == "scat":
        kw = dict(chi_scat=(cont * 0.3).astype(np.float32), dJnu=rng.uniform(-1e-2, 1e-2, cont.size).astype(np.float32))
    if mode == "suppress":
        kw = dict(suppress_masers=True)
    chi, S, nfl = O.work_to_transfer(xhat, ehat, beta, maser_chi_floor=1e-3, **kw)
    w1, w2 = d(xhat), d(ehat)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    gkw = {k: (d(v) if isinstance(v, np.ndarray) else v) for k, v in kw.items()}
    ctx.work_to_transfer(w1, w2, d(beta), maser_chi_floor=1e-3, n_floored=cnt, **gkw)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(w1.cpu().numpy(), chi)
    np.testing.assert_array_equal(w2.cpu().numpy(), S)
    assert int(cnt.item()) == nfl
    if mode != "suppress":
        assert nfl > 0


@pytest.mark.parametrize("ali", [False, True])
def test_background_emission_matches_oracle(ctx, ali):
    from test_oracle_background import fields as bfields
    rng, B, chi, sigma = bfields(n=200003, seed=9)
    Jfs = (B * rng.uniform(0.2, 2.0, B.size)).astype(np.float32)
    kw, gkw = {}, {}
    if ali:
        PsiJ = (rng.uniform(0.0, 1.5, B.size) / sigma).astype(np.float32)
        Jold = (B * rng.uniform(0.5, 3.0, B.size)).astype(np.float32)
        kw = dict(PsiJ=PsiJ, Jali_old=Jold, scat_op_max=0.8)
        gkw = dict(PsiJ=d(PsiJ), Jali_old=d(Jold), scat_op_max=0.8)
    ref = O.background_emission(Jfs, B, chi, sigma, wscat=0.75, **kw)
    gJ, geta, gdJ, gJali = d(Jfs), torch.zeros_like(d(B)), torch.zeros_like(d(B)), torch.zeros_like(d(B))
    nneg = torch.zeros(1, dtype=torch.int64, device="cuda")
    diag = torch.zeros(3, dtype=torch.float64, device="cuda")
    ctx.background_emission(gJ, d(B), d(chi), d(sigma), geta, dJnu=gdJ, Jali=gJali if ali else None, wscat=0.75,
                            n_neg=nneg, diag=diag, **gkw)
    torch.cuda.synchronize()
    # the FP64 ALI step rounds once to FP32: bit for bit; the FP32 source terms too
    np.testing.assert_array_equal(gJ.cpu().numpy(), ref["J"])
    np.testing.assert_array_equal(geta.cpu().numpy(), ref["eta"])
    np.testing.assert_array_equal(gdJ.cpu().numpy(), ref["dJnu"])
    if ali:
        np.testing.assert_array_equal(gJali.cpu().numpy(), ref["Jali"])
        assert int(nneg.item()) == ref["n_neg"] > 0
        np.testing.assert_allclose(diag.cpu().numpy(), [ref["op_raw"], ref["op_max"], ref["amp_max"]], rtol=1e-15)
