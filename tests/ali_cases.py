"""The oracle's ALI iteration: the same pipeline as dispatch_amd.ali.NlteProblem.rates
and .update, composed from the oracle's restatements (convert_to_work_units,
bb_opacity with the 'largest' tracking, work_to_transfer, short_char2 with split
rates, bb_rates 'largest' + preconditioning with Kahan sums, bf_rates,
normalize_rates, update_populations) on the problem's own inputs, copied to the
host.  Work cubes for the oracle's sweep are (nz+2, ny, nx, nf) with the ghost
planes; the problem's spectra are (nf, ncell), cells in (z, y, x) order."""
import numpy as np
import torch

from oracle import oracle as O


def h(t):
    return t.detach().cpu().numpy()


# ------------------------------------------------------------------ host context

n = lambda t: None if t is None else t.detach().cpu().numpy()


class OracleCtx:
    """The entry points NlteProblem calls, each by its oracle restatement on host
    tensors, in place: the driver's orchestration (windows, directions, Ng) runs on
    the CPU against the checker's arithmetic."""
    def sc_sweep(self, chi, S, gchi, gS, gI, chi500, chi500_up, g500, *, dir, ds, vec, layout, wj=0.0, J=None,
                 PsiJ=None, I_out=None, Psi_out=None, Irem_out=None, escape_out=None, exit_I=None,
                 split_rates=False, get_rates=None):
        assert layout == "fmajor"
        nf, nz, ny, nx = chi.shape
        def work(c, g):
            cube = np.moveaxis(n(c), 0, 3)           # (nz, ny, nx, nf)
            gp = np.moveaxis(n(g), 0, 2)[None]       # (1, ny, nx, nf)
            z = np.zeros_like(gp)
            return np.ascontiguousarray(np.concatenate([gp, cube, z] if dir > 0 else [z, cube, gp]), np.float32)
        I, psi, irem, esc = O.short_char2(work(chi, gchi), work(S, gS), work(torch.zeros_like(chi), gI), n(chi500),
                                          n(chi500_up), n(g500), dir, ds, vec, split_rates=split_rates,
                                          get_rates=True)
        back = lambda a: torch.from_numpy(np.ascontiguousarray(np.moveaxis(a, 3, 0)))
        Ii = back(I[1:-1])
        if J is not None: J.add_(float(np.float32(wj)) * Ii.to(J.dtype))
        if PsiJ is not None: PsiJ.add_(float(np.float32(wj)) * back(psi).to(PsiJ.dtype))
        if I_out is not None: I_out.copy_(Ii)
        if Psi_out is not None: Psi_out.copy_(back(psi))
        if Irem_out is not None: Irem_out.copy_(back(irem))
        if escape_out is not None: escape_out.copy_(back(esc))

    def lte_background(self, nu, temp, beta, Bnu):
        b, B = O.lte_background(n(nu), n(temp))
        beta.copy_(torch.from_numpy(b.reshape(beta.shape)))
        Bnu.copy_(torch.from_numpy(B.reshape(Bnu.shape)))

    def convert_to_work_units(self, chi_c, xhat, ehat, beta):
        x, e = O.convert_to_work_units(n(chi_c), n(ehat), n(beta))
        xhat.copy_(torch.from_numpy(x)); ehat.copy_(torch.from_numpy(e))

    def bb_opacity(self, d, xhat, ehat, wmu=0.0, split=None, split_offset=0):
        dn = {k: n(v) for k, v in d.items()}
        sp = {k: n(v) for k, v in split.items()}
        r = O.bb_opacity(dn, n(xhat), n(ehat), wmu=wmu, split=sp, split_offset=split_offset)
        xhat.copy_(torch.from_numpy(r["xhat"])); ehat.copy_(torch.from_numpy(r["ehat"]))
        for k in split:
            split[k].copy_(torch.from_numpy(np.asarray(r[k])))

    def window_opacity(self, d, chi_cont, Bnu, beta, xhat, ehat, split=None, split_offset=0):
        # the fused entry point's definition: the sequence it replaces
        xhat.zero_()
        torch.mul(chi_cont, Bnu, out=ehat)
        self.convert_to_work_units(chi_cont, xhat, ehat, beta)
        split["split_opac1"].copy_(ehat)
        split["split_chi1"].copy_(xhat)
        for k in ("split_opac2", "split_chi2", "split_id", "split_chi_id"):
            split[k].zero_()
        self.bb_opacity(d, xhat, ehat, split=split, split_offset=split_offset)
        self.work_to_transfer(xhat, ehat, beta, split_opac1=split["split_opac1"], split_opac2=split["split_opac2"])

    def work_to_transfer(self, xhat, ehat, beta, split_opac1=None, split_opac2=None):
        s1, s2 = np.ascontiguousarray(n(split_opac1)), np.ascontiguousarray(n(split_opac2))
        c, s, _ = O.work_to_transfer(n(xhat), n(ehat), n(beta), split_opac1=s1, split_opac2=s2)
        xhat.copy_(torch.from_numpy(c)); ehat.copy_(torch.from_numpy(s))
        split_opac1.copy_(torch.from_numpy(s1)); split_opac2.copy_(torch.from_numpy(s2))

    def bb_rates(self, d, wmu, wtot, gij, gji, precond=False, own_norm=False, largest=False, split=None,
                 split_offset=0):
        out = O.bb_rates({k: n(v) for k, v in d.items()}, wmu, wtot, n(gij), n(gji), precond=precond,
                         own_norm=own_norm, largest=largest, kahan=True,
                         split={k: n(v) for k, v in split.items()}, split_offset=split_offset)
        gij.copy_(torch.from_numpy(out[0])); gji.copy_(torch.from_numpy(out[1]))

    def bf_rates(self, d, nline, gij, gji, precond=False):
        a, b = O.bf_rates({k: n(v) for k, v in d.items()}, nline, n(gij), n(gji), precond=precond)
        gij.copy_(torch.from_numpy(a)); gji.copy_(torch.from_numpy(b))

    def normalize_rates(self, atom, nstar, gij, gji, wphi=None):
        a, b = O.normalize_rates(atom, n(nstar), n(gij), n(gji))
        gij.copy_(torch.from_numpy(a)); gji.copy_(torch.from_numpy(b))

    def rate_solve(self, atom, e_kT, temp, totn, nstar, b, gij, gji, cup, **params):
        bn, res = O.update_populations(atom, n(e_kT), n(temp), n(totn), n(nstar), n(b), n(gij), n(gji), n(cup),
                                       **params)
        self.last = dict(gij=n(gij), gji=n(gji), b_in=n(b).copy())
        if res["status"] != 0:
            raise RuntimeError(f"rate solve status {res['status']} epos {res['epos']}")
        b.copy_(torch.from_numpy(bn))
        return res



def to_work(a, p, ghost_lo, ghost_hi):
    """(nf, ncell) -> (nz+2, ny, nx, nf) with the two ghost planes (nf, ny*nx)."""
    c = p.cfg
    nf = a.shape[0]
    cube = np.moveaxis(a.reshape(nf, c.nz, c.ny, c.nx), 0, 3)
    lo = np.moveaxis(ghost_lo.reshape(nf, c.ny, c.nx), 0, 2)[None]
    hi = np.moveaxis(ghost_hi.reshape(nf, c.ny, c.nx), 0, 2)[None]
    return np.ascontiguousarray(np.concatenate([lo, cube, hi]), np.float32)


def from_work(a):
    """(nz, ny, nx, nf) -> (nf, ncell)."""
    nz, ny, nx, nf = a.shape
    return np.ascontiguousarray(np.moveaxis(a, 3, 0).reshape(nf, -1))


def oracle_sweep(p, r, chi, S, ghost_B):
    """The problem's direction r by the oracle's short_char2 with split rates:
    returns (I, Psi, Irem, escape) as (nf, ncell)."""
    c = p.cfg
    nf = chi.shape[0]
    v = p.vec[r]
    up = v[2] > 0
    r500 = h(p.r500)
    zero = np.zeros((nf, c.nx * c.ny), np.float32)
    if up:
        gchi = (chi[:, :c.nx * c.ny].astype(np.float64) * r500[None]).astype(np.float32)
        gB = h(ghost_B)
        wc, ws, wi = to_work(chi, p, gchi, zero), to_work(S, p, gB, zero), to_work(np.zeros_like(S), p, gB, zero)
        g500 = h(p.box.g500_bot)
    else:
        wc, ws, wi = to_work(chi, p, zero, zero), to_work(S, p, zero, zero), to_work(np.zeros_like(S), p, zero, zero)
        g500 = h(p.box.g500_top)
    I, psi, irem, esc = O.short_char2(wc, ws, wi, h(p.box.chi500), h(p.cu[r]), g500, 1 if up else -1, p.box.ds, v,
                                      split_rates=True, get_rates=True)
    return from_work(I[1:-1]), from_work(psi), from_work(irem), from_work(esc)


def oracle_rates(p):
    """gamma_ij, gamma_ji (nrad, ncell) of one iteration at the problem's current b."""
    atom = p.atom
    nrad = len(atom.ti)
    gij = np.zeros((nrad, p.ncell), np.float32)
    gji = np.zeros_like(gij)
    b, nstar = h(p.b), h(p.nstar)
    for w in p.windows:
        l0, l1 = w.lines[0], w.lines[-1] + 1
        const = {k: h(t) for k, t in w.const.items()}
        chi_c, beta, Bnu = h(w.chi_c), h(w.beta), h(w.Bnu)
        nf = w.nf
        for r, v in enumerate(p.vec):
            vproj = h((p.vz * (v[2] / 2.99792458e10)).float())
            d = dict(const, b=b, nstar=nstar, pop_ratio=h(p.pop_ratio), inv_df=h(p.inv_df), vproj=vproj,
                     e_kT=h(p.e_kT32), adamp=h(p.adamp[l0:l1]))
            ehat = (chi_c * Bnu).astype(np.float32)
            xhat, ehat = O.convert_to_work_units(chi_c, ehat, beta)
            z = np.zeros((nf, p.ncell), np.float32)
            zi = np.zeros((nf, p.ncell), np.int32)
            # the continuum seeds the split tracking as contribution 0
            sp0 = dict(split_opac1=ehat.copy(), split_opac2=z, split_chi1=xhat.copy(), split_chi2=z.copy(),
                       split_id=zi, split_chi_id=zi.copy())
            op = O.bb_opacity(d, xhat, ehat, wmu=float(p.wj[r]), split=sp0, split_offset=0)
            so1 = np.ascontiguousarray(op["split_opac1"], np.float32)
            so2 = np.ascontiguousarray(op["split_opac2"], np.float32)
            chi, S, _ = O.work_to_transfer(op["xhat"], op["ehat"], beta, split_opac1=so1, split_opac2=so2)
            I, Psi, Irem, esc = oracle_sweep(p, r, chi, S, w.ghost_B)
            sp = dict(Irem=Irem, escape=esc, chi=chi, split_opac1=so1, split_opac2=so2, split_chi1=op["split_chi1"],
                      split_chi2=op["split_chi2"], split_id=op["split_id"], split_chi_id=op["split_chi_id"])
            out = O.bb_rates(dict(d, Inu=I, Psi=Psi, Bnu=Bnu, beta=beta), float(p.wj[r]), 1.0, gij[l0:l1],
                             gji[l0:l1], precond=True, own_norm=True, largest=True, kahan=True, split=sp,
                             split_offset=0)
            gij[l0:l1], gji[l0:l1] = out[0], out[1]
    bf = {k: h(t) for k, t in p.bf.items() if isinstance(t, torch.Tensor)}
    bf["b"] = b
    # the continuum windows' J and PsiJ by the oracle's sweeps
    cchi, cB = h(p.cont["chi_c"]), h(p.cont["B"])
    J = np.zeros_like(cchi)
    PsiJ = np.zeros_like(cchi)
    for r in range(len(p.vec)):
        I, Psi, _, _ = oracle_sweep(p, r, cchi, cB, p.cont["ghost_B"])
        J = (J + np.float32(p.wj[r]) * I).astype(np.float32)
        PsiJ = (PsiJ + np.float32(p.wj[r]) * Psi).astype(np.float32)
    bf["Jnu"], bf["PsiJ"] = J, PsiJ
    gij, gji = O.bf_rates(bf, atom.nline, gij, gji, precond=False)
    return O.normalize_rates(atom, nstar, gij, gji)


def oracle_update(p, gij, gji):
    return O.update_populations(p.atom, h(p.e_kT64), h(p.temp), h(p.totn), h(p.nstar), h(p.b), gij, gji, h(p.cup))
