"""GPU parity of the LTE background factors and the boundary planes
(m3drt_lte_background, m3drt_set_upstream_boundary, m3drt_set_downstream_boundary)
against the oracle (SRC/rt_patch_mod.f90:2004-2023, 3270-3374, 3471-3511):

- chi of the LTE ghost, the PCHIP ghost (with its clamp at 0) and the copies are
  bit for bit;
- beta takes the FP64 exponential rounded once, the reference the FP32 one: they
  agree to one ulp, and Bnu is beta/(1 - beta) of the kernel's own beta bit for bit
  (so both agree with the oracle bit for bit wherever beta does, which is nearly
  everywhere).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def d(a, dt=None):
    return torch.from_numpy(np.ascontiguousarray(a if dt is None else a.astype(dt))).cuda()


def ulp_diff(a, b):
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


def check_beta_B(beta, B, beta_o, B_o):
    assert ulp_diff(beta, beta_o).max() <= 1
    assert np.array_equal(B, (beta / (np.float32(1) - beta)).astype(np.float32))
    same = beta == beta_o
    assert same.mean() > 0.99 and np.array_equal(B[same], B_o[same])


def test_lte_background(ctx):
    rng = np.random.default_rng(11)
    nu = np.concatenate([np.geomspace(1e12, 3e16, 500), [1e17]])
    T = np.concatenate([rng.uniform(2500, 30000, 3000), [3000.0]])
    beta, B = torch.empty((nu.size, T.size), dtype=torch.float32, device="cuda"), None
    B = torch.empty_like(beta)
    ctx.lte_background(d(nu), d(T), beta, B)
    torch.cuda.synchronize()
    bo, Bo = O.lte_background(nu, T)
    check_beta_B(beta.cpu().numpy(), B.cpu().numpy(), bo, Bo)
    assert np.any(bo == np.finfo(np.float32).tiny)


def planes(rng, nf, ny, nx):
    return (10 ** rng.uniform(-9, -4, (nf, ny, nx))).astype(np.float32)


@pytest.mark.parametrize("shape", [(7, 5, 9), (64, 120, 120)])
def test_upstream_boundary_kinds(ctx, shape):
    nf, ny, nx = shape
    rng = np.random.default_rng(nf)
    chi1, chi2, S1, S2 = (planes(rng, nf, ny, nx) for _ in range(4))
    # a few cells where the extrapolation goes negative (clamped) and flat pairs
    chi2.ravel()[::7] = chi1.ravel()[::7] * 1.6
    S2.ravel()[::11] = S1.ravel()[::11]
    ghost = [torch.full((nf, ny, nx), -1.0, device="cuda") for _ in range(3)]
    # top: zeros
    ctx.set_upstream_boundary("top", ghost)
    torch.cuda.synchronize()
    assert all(float(g.abs().max()) == 0.0 for g in ghost)
    # PCHIP
    ds3 = 10e5
    ctx.set_upstream_boundary("pchip", ghost, chi1=d(chi1), S1=d(S1), chi2=d(chi2), S2=d(S2), ds3=ds3)
    torch.cuda.synchronize()
    ec = np.maximum(O.pchip_ghost(chi1, chi2, chi2, ds3), 0).reshape(chi1.shape)
    es = np.maximum(O.pchip_ghost(S1, S2, S2, ds3), 0).reshape(chi1.shape)
    assert np.array_equal(ghost[0].cpu().numpy(), ec) and np.any(ec == 0)
    assert np.array_equal(ghost[1].cpu().numpy(), es)
    assert np.array_equal(ghost[2].cpu().numpy(), es)
    # LTE
    nu = np.linspace(5.99e14, 6.0e14, nf)
    g500 = (10 ** rng.uniform(-7, -6, (ny, nx))).astype(np.float32)
    c500 = (10 ** rng.uniform(-7.3, -6.3, (ny, nx))).astype(np.float32)
    T = rng.uniform(6000, 9000, (ny, nx)).astype(np.float32)
    ctx.set_upstream_boundary("lte", ghost, chi1=d(chi1), g500=d(g500), chi500_1=d(c500), T_ghost=d(T), nu=d(nu))
    torch.cuda.synchronize()
    oc, oS, oI = O.lte_ghost(nu, chi1.reshape(nf, -1), g500, c500, T)
    assert np.array_equal(ghost[0].cpu().numpy().reshape(nf, -1), oc)
    gS = ghost[1].cpu().numpy().reshape(nf, -1)
    bo, Bo = O.lte_background(nu, T.ravel().astype(np.float64))
    assert ulp_diff(gS, oS).max() <= 8 and np.mean(gS == oS) > 0.99
    assert np.array_equal(ghost[2].cpu().numpy().reshape(nf, -1), gS)
    # upstream slab's exit plane, and the downstream copy back
    mem = [d(planes(rng, nf, ny, nx)) for _ in range(3)]
    ctx.set_upstream_boundary("upstream", ghost, mem=mem)
    out = [torch.zeros_like(m) for m in mem]
    ctx.set_downstream_boundary(mem[0], mem[1], mem[2], out)
    torch.cuda.synchronize()
    for g, m, o in zip(ghost, mem, out):
        assert torch.equal(g, m) and torch.equal(o, m)


def test_missing_inputs_are_errors(ctx):
    from dispatch_amd import m3drt
    ghost = [torch.zeros((2, 4, 4), device="cuda") for _ in range(3)]
    with pytest.raises(m3drt.M3drtError, match="LTE ghost needs"):
        ctx.set_upstream_boundary("lte", ghost)
    with pytest.raises(m3drt.M3drtError, match="PCHIP"):
        ctx.set_upstream_boundary("pchip", ghost, chi1=ghost[0], S1=ghost[1], chi2=ghost[0], S2=ghost[1])
