"""Pins the oracle's dense restatement of update_eta_mask
(SRC/rt_patch_mod.f90:3671-3802) against a line-by-line Python restatement of the
reference's own sparse algorithm (rows appended in (z, y, x) loop order by
cell_row, eta_src/eta_mask indexed by row, the ghost plane's eta_gsrc), over
several successive calls so the flags' stickiness is exercised.  The flag sets
must agree exactly; the row numbering is the reference's internal bookkeeping.
"""
import numpy as np
import pytest

from oracle import oracle as O

Q = np.float32(1e-2)


def fire_of(chi, S, beta):
    chi, S = np.float32(chi), np.float32(S)
    one_m_b = np.float32(1.0) - np.float32(beta) if beta is not None else np.float32(1.0)
    t = np.float32(abs(chi)) - Q * (chi + one_m_b * chi * S)
    return min(chi, t) <= 0.0


class SparseEta:
    """The reference's data structure and loops, in plain Python (small cases)."""

    def __init__(self, nf, nx, ny, nz, dir, rx, ry):
        self.nf, self.nx, self.ny, self.nz, self.dir, self.rx, self.ry = nf, nx, ny, nz, dir, rx, ry
        self.eta_row = np.zeros((nz, ny, nx), np.int64)
        self.src, self.mask = [], []           # rows of length nf
        self.gsrc = np.zeros((nf, ny, nx), bool)

    def cell_row(self, x, y, z):
        if self.eta_row[z, y, x] == 0:
            self.src.append(np.zeros(self.nf, bool))
            self.mask.append(np.zeros(self.nf, bool))
            self.eta_row[z, y, x] = len(self.src)
        return self.eta_row[z, y, x]

    def spread(self, x, y, z, new):
        for j in range(-self.ry, self.ry + 1):
            for i in range(-self.rx, self.rx + 1):
                r = self.cell_row((x + i) % self.nx, (y + j) % self.ny, z)
                self.mask[r - 1] |= new

    def update(self, chi, S, beta, gchi, gS, ghost):
        nf, nx, ny, nz = self.nf, self.nx, self.ny, self.nz
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    fire = np.array([fire_of(chi[z, f, y, x], S[z, f, y, x], beta[z, f, y, x]) for f in range(nf)])
                    if not fire.any():
                        continue
                    row = self.eta_row[z, y, x]
                    new = fire & ~self.src[row - 1] if row > 0 else fire
                    if not new.any():
                        continue
                    row = self.cell_row(x, y, z)
                    self.src[row - 1] |= new
                    self.spread(x, y, z, new)
                    zn = z + self.dir
                    if 0 <= zn < nz:
                        self.spread(x, y, zn, new)
        if ghost:
            z1 = 0 if self.dir > 0 else nz - 1
            for y in range(ny):
                for x in range(nx):
                    fire = np.array([fire_of(gchi[f, y, x], gS[f, y, x], None) for f in range(nf)])
                    new = fire & ~self.gsrc[:, y, x]
                    if not new.any():
                        continue
                    self.gsrc[:, y, x] |= new
                    self.spread(x, y, z1, new)

    def dense(self):
        src = np.zeros((self.nz, self.nf, self.ny, self.nx), np.uint8)
        mask = np.zeros_like(src)
        for z, y, x in zip(*np.nonzero(self.eta_row)):
            r = self.eta_row[z, y, x] - 1
            src[z, :, y, x] = self.src[r]
            mask[z, :, y, x] = self.mask[r]
        return src, mask, self.gsrc.astype(np.uint8)


def fields(rng, nf, nx, ny, nz, frac):
    chi = (10 ** rng.uniform(-6, -4, (nz, nf, ny, nx))).astype(np.float32)
    S = rng.uniform(0.01, 1.0, chi.shape).astype(np.float32)
    beta = rng.uniform(0.1, 0.9, chi.shape).astype(np.float32)
    neg = rng.uniform(size=chi.shape) < frac
    chi = np.where(neg, -0.5 * chi, chi).astype(np.float32)
    # near-transparent pairs: |chi| just above and below the q threshold
    near = rng.uniform(size=chi.shape) < frac
    S = np.where(near, (np.abs(chi) / Q - chi) / ((1 - beta) * chi) * rng.uniform(0.9, 1.1, chi.shape), S)
    gchi = (10 ** rng.uniform(-6, -4, (nf, ny, nx))).astype(np.float32)
    gchi = np.where(rng.uniform(size=gchi.shape) < frac, -gchi, gchi).astype(np.float32)
    gS = rng.uniform(0.01, 1.0, gchi.shape).astype(np.float32)
    return chi, S.astype(np.float32), beta, gchi, gS


@pytest.mark.parametrize("dir,rx,ry,ghost", [(1, 1, 2, True), (-1, 2, 1, False), (1, 0, 0, True)])
def test_dense_equals_reference_sparse(dir, rx, ry, ghost):
    rng = np.random.default_rng(rx * 10 + ry + (dir > 0))
    nf, nx, ny, nz = 3, 6, 5, 4
    ref = SparseEta(nf, nx, ny, nz, dir, rx, ry)
    src = np.zeros((nz, nf, ny, nx), np.uint8)
    mask = np.zeros_like(src)
    gsrc = np.zeros((nf, ny, nx), np.uint8)
    for call in range(3):  # flags are sticky across calls
        chi, S, beta, gchi, gS = fields(rng, nf, nx, ny, nz, 0.03)
        ref.update(chi, S, beta, gchi, gS, ghost)
        src, mask, gsrc = O.update_eta_mask(chi, S, beta, gchi, gS, src, mask, gsrc, dir, rx, ry, ghost)
        rs, rm, rg = ref.dense()
        np.testing.assert_array_equal(src, rs)
        np.testing.assert_array_equal(mask, rm)
        np.testing.assert_array_equal(gsrc, rg)
    assert mask.sum() > src.sum() > 0  # the cases fire and spread
