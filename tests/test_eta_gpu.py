"""GPU parity of update_eta_mask on dense flags (m3drt_update_eta_mask) against the
oracle over three successive calls (the flags are sticky): src, mask and the ghost
plane's flags bit for bit."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from test_oracle_eta import fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dir,rx,ry,ghost", [(1, 2, 1, True), (-1, 1, 3, False)])
def test_eta_mask_matches_oracle(ctx, dir, rx, ry, ghost):
    rng = np.random.default_rng(7 + rx)
    nf, nx, ny, nz = 20, 33, 17, 9
    src = np.zeros((nz, nf, ny, nx), np.uint8)
    mask = np.zeros_like(src)
    gsrc = np.zeros((nf, ny, nx), np.uint8)
    g = {k: torch.from_numpy(v.copy()).cuda() for k, v in dict(src=src, mask=mask, gsrc=gsrc).items()}
    for call in range(3):
        chi, S, beta, gchi, gS = fields(rng, nf, nx, ny, nz, 0.01)
        src, mask, gsrc = O.update_eta_mask(chi, S, beta, gchi, gS, src, mask, gsrc, dir, rx, ry, ghost)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        ctx.update_eta_mask(t(chi), t(S), t(beta), g["src"], g["mask"], dir=dir, rx=rx, ry=ry,
                            ghost_chi=t(gchi) if ghost else None, ghost_S=t(gS) if ghost else None,
                            ghost_src=g["gsrc"] if ghost else None)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(g["src"].cpu().numpy(), src)
        np.testing.assert_array_equal(g["mask"].cpu().numpy(), mask)
        if ghost:
            np.testing.assert_array_equal(g["gsrc"].cpu().numpy(), gsrc)
    assert mask.sum() > src.sum() > 0
