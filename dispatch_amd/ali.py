"""3D NLTE ALI iterations on one GPU: the per-iteration pipeline of OpenMULTI3D's
rt_patch_t%execute and nlte_patch (SRC/rt_patch_mod.f90:960-1274,
SRC/nlte_patch_mod.f90:561-838) for one model atom, in the reference's default
formulation (rate_formulation='largest' with preconditioning: split rates).

Per iteration, per frequency window of lines, per direction:
  1. work units: the LTE continuum (chi_c, eta_c = chi_c B) converted
     (m3drt_convert_to_work_units, SRC/rt_patch_mod.f90:3518-3536), then every line
     of the window added with the 'largest' tracking (m3drt_bb_opacity:
     add_bound_bound_opacities + add_largest, :2689-2831, 3150-3169);
  2. work_to_transfer_units (m3drt_work_to_transfer, :3571-3651) -> chi, S;
  3. the formal solution of the direction (m3drt_sc_sweep, short_char2) with the
     intensity, Psi, Irem and escape of every (f, cell);
  4. the line-rate integrals of the direction (m3drt_bb_rates, bound_bound_rates
     :2852-3142, Kahan sums).
Then the continuum rates from the continuum windows' J (m3drt_bf_rates,
:3193-3259), normalize_profiles (m3drt_normalize_rates) and the FP64
statistical-equilibrium update of every cell (m3drt_rate_solve_f64: rate matrices,
batched LU without pivoting, limiter, max |dn/n|).

Spectra are (f, cell) with the cells flat in (z, y, x) order; the sweep takes the
same memory as (nf, nz, ny, nx) ("fmajor"), so the opacity, transfer and rate
kernels share one layout with no transposes.  A window's arrays are allocated once
and reused over directions, windows and iterations.

The workload: a 60 x 60 x 120 box from the 3D LTE generator, the synthetic
21-level atom (20 bound levels, continuum; 80 lines, 20 continua), 120 frequencies
per line, van Regemorter collisions, 16 rays (2 Gauss mu x 4 azimuths, mirrored);
iterate to max |dn/n| < 1e-4 (the command line caps the count at 60).  The line-of-sight velocity is the
generator's vertical velocity projected on each ray.  The continuum windows carry
the background alone (the atom's own bound-free opacity is not added to the
transfer), so their J is computed once.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import m3drt, synth
from .constants import BK, CC, E_K, H_4PI, TWOH_C2, UU
PI_E2_MC = 0.0265400   # pi e^2 / (m_e c) [cm^2 Hz]


def _trapezoid(x):
    w = np.zeros_like(x)
    d = np.diff(x)
    w[:-1] += 0.5 * d
    w[1:] += 0.5 * d
    return w


@dataclass
class AliConfig:
    nx: int = 60
    ny: int = 60
    nz: int = 120
    nq: int = 120             # frequencies per line
    qmax: float = 4.0e-5      # half-width of a line window in (nu - nu0) / nu0
    ncont_f: int = 24         # frequencies per continuum
    lines_per_window: int = 20
    nphi: int = 4             # azimuths per Gauss mu: 2 x 4 x 2 = 16 rays
    nbound: int = 20          # bound levels; every one has a continuum when ncont = nbound
    nline: int = 80
    ncont: int = 20
    atom_seed: int = 20261204
    box_seed: int = 20261204
    ne_scale: float = 1.0     # multiplies the electron density (collisional coupling)
    max_iter: int = 40
    conv: float = 1e-4
    ng: bool = True           # Ng acceleration (the reference's automatic trigger)
    ng_step: int = 0          # > 0: the reference's fixed schedule instead (&nlte_params ng_step,
    ng_start: int = 8         # ng_start; SRC/atom_mod.f90:278-287, 2500-2502); 1 or 2 is raised
                              # to 3 (SRC/m3d_patch_mod.f90:472-476)
    sor_w: float = 1.0        # &nlte_params sor_w: the update factor e -> e**sor_w
                              # (SRC/atom_mod.f90:255-266, 604-610); 1 is undamped
    conv_pop_floor: float = 0.0  # &nlte_params conv_pop_floor: a level holding less than this share
                              # of the atoms counts in the convergence measure with its change
                              # scaled by its share over the floor (SRC/atom_mod.f90:233-236,
                              # 673-686); 0 weights every level fully
    # the automatic Ng gate's &nlte_params (SRC/m3d_patch_mod.f90:405-406): overrides of
    # NgAcceleration's defaults by name (rate_min, rate_tol, clamp_max, min_iters_left,
    # cond_min, max_change, min_gap)
    ng_gate: dict = field(default_factory=dict)

    def __post_init__(self):
        # Ng needs three iterates after an extrapolation before the next one
        if 0 < self.ng_step < 3:
            import warnings
            warnings.warn(f"NG acceleration needs 3 initial iterations; ng_step {self.ng_step} set to 3")
            self.ng_step = 3
    precond: bool = True      # preconditioned (ALI) rates; False: plain Lambda iteration
    fused: bool = True        # one kernel per (window, direction) for continuum, lines and units
    graph: bool = False       # replay the rate pass of an iteration as one captured HIP graph


@dataclass
class _Window:
    lines: list
    nf: int
    chi_c: torch.Tensor
    beta: torch.Tensor
    Bnu: torch.Tensor
    ghost_B: torch.Tensor
    const: dict = field(default_factory=dict)


def ng_fixed_fires(it: int, step: int, start: int) -> bool:
    """The reference's fixed Ng schedule, mod(iter - ng_start, ng_step) == 1
    (SRC/atom_mod.f90:2500-2502), with Fortran's mod: the result takes the dividend's
    sign, so nothing fires before ng_start."""
    d = it - start
    return d - step * int(d / step) == 1


class NgAcceleration:
    """Ng's three-difference extrapolation in ln b (SRC/ng_mod.f90, applied as
    atom_type%ng_acceleration does, SRC/atom_mod.f90:2469-2554): the Gram matrix of
    the differences of the last four iterates is accumulated each iteration and
    promoted at the next, whose gate fires only in the iteration's asymptotic
    regime (stationary contraction rate between 0.5 and 1, few limited corrections,
    a well-conditioned fit, headroom left); the extrapolated b is renormalised to
    the total density and the step vetoed if it moves a level by more than a factor
    11.  The gate's constants are the reference's &nlte_params defaults."""
    rate_min, rate_tol, clamp_max, min_iters_left = 0.5, 0.15, 1e-3, 3
    cond_min, max_change, min_gap = 1e-6, 10.0, 4
    GATE = ("rate_min", "rate_tol", "clamp_max", "min_iters_left", "cond_min", "max_change", "min_gap")

    def __init__(self, **gate):
        for k, v in gate.items():
            if k not in self.GATE:
                raise ValueError(f"unknown Ng gate parameter {k!r} (one of {', '.join(self.GATE)})")
            setattr(self, k, type(getattr(self, k))(v))
        self.d_cur = np.zeros((3, 3))
        self.d_done = np.zeros((3, 3))
        self.s_cur = self.s_done = 0.0
        self.s_prev = -1.0
        self.r = self.r_prev = -1.0
        self.nclamp_cur = self.ntot_cur = self.nclamp_done = self.ntot_done = 0
        self.iter_fired = -(1 << 30)

    def promote(self):
        self.d_done, self.s_done = self.d_cur, self.s_cur
        self.nclamp_done, self.ntot_done = self.nclamp_cur, self.ntot_cur
        self.d_cur, self.s_cur = np.zeros((3, 3)), 0.0
        self.nclamp_cur = self.ntot_cur = 0
        if self.s_done > 0:
            self.r_prev = self.r
            self.r = math.sqrt(self.s_done / self.s_prev) if self.s_prev > 0 else -1.0
            self.s_prev = self.s_done

    def condition(self):
        d = self.d_done
        scale = d[1, 1] * d[2, 2]
        return (d[1, 1] * d[2, 2] - d[1, 2] * d[2, 1]) / scale if scale > 0 else 0.0

    def should_fire(self, it, err, convlim):
        if self.s_done <= 0 or self.s_prev <= 0:
            return False
        if it - self.iter_fired < self.min_gap or self.ntot_done <= 0:
            return False
        if self.nclamp_done / self.ntot_done > self.clamp_max:
            return False
        if self.r <= 0 or self.r_prev <= 0 or abs(self.r - self.r_prev) > self.rate_tol * self.r:
            return False
        if self.r < self.rate_min or self.r >= 1.0:
            return False
        if abs(err) < convlim * self.r ** (-self.min_iters_left):
            return False
        return self.condition() > self.cond_min

    def accumulate(self, b, oldb, nclamp):
        n1, n2, n3 = (torch.log(o) for o in oldb)
        n4 = torch.log(b)
        d0 = n4 - n3
        d1 = d0 - (n3 - n2)
        d2 = d0 - (n2 - n1)
        dot = lambda u, v: float((u * v).sum())
        self.d_cur[0, 1] += dot(d0, d1)
        self.d_cur[0, 2] += dot(d0, d2)
        self.d_cur[1, 1] += dot(d1, d1)
        self.d_cur[1, 2] += dot(d1, d2)
        self.d_cur[2, 1] = self.d_cur[1, 2]
        self.d_cur[2, 2] += dot(d2, d2)
        self.s_cur += dot(d0, d0)
        self.nclamp_cur += int(nclamp)
        self.ntot_cur += b.numel()

    def extrapolate(self, b, oldb, nstar, totn):
        """Apply the step to b in place; returns the largest relative change, or None
        when vetoed."""
        d = self.d_done
        den = d[1, 1] * d[2, 2] - d[1, 2] * d[2, 1]
        a = (d[0, 1] * d[2, 2] - d[0, 2] * d[2, 1]) / den
        bb = (d[0, 2] * d[1, 1] - d[0, 1] * d[1, 2]) / den
        n2, n3, n4 = torch.log(oldb[1]), torch.log(oldb[2]), torch.log(b)
        e = (1.0 - a - bb) * n4 + a * n3 + bb * n2
        worst = float((e - n4).abs().max())
        if not (math.exp(worst) - 1.0 <= self.max_change and math.isfinite(a) and math.isfinite(bb)):
            return None
        bn = torch.exp(e)
        bn *= totn[None, :] / (bn * nstar).sum(dim=0, keepdim=True)
        change = float((bn / b - 1.0).abs().max())
        b.copy_(bn)
        return change


class NlteProblem:
    def __init__(self, ctx: m3drt.Context, cfg: AliConfig = AliConfig(), sweep_dtype=torch.float32,
                 device="cuda"):
        """ctx: the m3drt context whose entry points every step calls; device: where
        the arrays live (the library's device; another one only for a context object
        with the same methods)."""
        self.ctx, self.cfg, self.sweep_dtype, self.device = ctx, cfg, sweep_dtype, device
        box = synth.make_box(cfg.nx, cfg.ny, cfg.nz, nf=2, seed=cfg.box_seed, device=device)
        self.box = box
        self.ncell = ncell = cfg.nx * cfg.ny * cfg.nz
        dev = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=self.device)
        T = box.T.cpu().numpy().reshape(-1)
        c5 = box.chi500.double().cpu().numpy().reshape(-1)
        # densities from the tau500 scale: n_H near 1e17 cm^-3 where tau500 = 1, and
        # electrons from the metals below ~6000 K, rising with T above
        H = (cfg.nz - 1) * box.ds[2] / (7.0 * math.log(10.0))
        nH = 1e17 * c5 * H
        ne = cfg.ne_scale * 1e-4 * nH * np.minimum((T / 5800.0) ** 6, 30.0)
        totn = 3e-5 * nH
        atom = synth.make_atom(nbound=cfg.nbound, nline=cfg.nline, ncont=cfg.ncont, seed=cfg.atom_seed)
        self.atom = atom
        nstar = synth.lte_populations(atom, T, ne, totn)
        self.T = T
        self.nstar = dev(nstar, torch.float64)
        self.totn = dev(totn, torch.float64)
        self.b = torch.ones_like(self.nstar)
        self.e_kT64 = dev(E_K / T, torch.float64)
        self.temp = dev(T, torch.float64)
        # Seaton's ionisation with gbar = 0.1 and the edge cross section
        self.cup = dev(synth.van_regemorter(atom, T, ne, nstar, ion_scale=0.1 * float(atom.sigma0[0])),
                       torch.float64)
        self.atom_dev = m3drt.AtomOnDevice(atom) if device == "cuda" else atom
        # per cell: Doppler width (mass 56 plus 1 km/s), damping, e/kT
        vth = np.sqrt(2 * BK * T / (56.0 * UU) + 1e5 ** 2) / CC
        self.inv_df = dev(1.0 / vth, torch.float32)
        self.e_kT32 = dev(E_K / T, torch.float32)
        self.adamp = dev(np.tile((2e-3 + 1e-2 * (ne / 1e14) ** 0.3)[None, :], (atom.nline, 1)), torch.float32)
        self.pop_ratio = torch.ones(ncell, dtype=torch.float64, device=self.device)
        self.vz = box.vz.reshape(-1)
        self.vec, self.wj = synth.gauss_rays(cfg.nphi)
        self.cu = [synth.chi500_upwind(box, v) for v in self.vec]
        self.r500 = (box.g500_bot.double() / box.chi500[0].double()).reshape(-1)
        self.windows = [self._window(lines) for lines in
                        [list(range(w, min(atom.nline, w + cfg.lines_per_window)))
                         for w in range(0, atom.nline, cfg.lines_per_window)]]
        self._graph = None
        self._graph_out = None
        self.history = []
        self.ng_history = []
        self.ng = NgAcceleration(**cfg.ng_gate) if cfg.ng else None
        self._oldb = []
        self._iter = 0
        self._continuum()

    # ---------------------------------------------------------------- spectra
    def _background(self, nu):
        """LTE background continuum on frequencies nu: chi_c, beta = e^(-h nu/kT) and
        B = beta / (1 - beta), FP32 (nf, ncell), from m3drt_lte_background
        (lte_background_opacities' factors); and B of the ghost plane below."""
        c5 = self.box.chi500.double().reshape(1, -1)
        nu = torch.as_tensor(nu, dtype=torch.float64, device=self.device)
        chi_c = c5 * (0.7 + 0.3 * (6e14 / nu[:, None]) ** 3)
        planes = []
        for T in (self.temp, self.box.T_ghost_bot.reshape(-1).double().contiguous()):
            beta = torch.empty((nu.numel(), T.numel()), dtype=torch.float32, device=self.device)
            B = torch.empty_like(beta)
            self.ctx.lte_background(nu, T.contiguous(), beta, B)
            planes.append((beta, B))
        return chi_c.float(), planes[0][0], planes[0][1], planes[1][1]

    def _window(self, lines):
        """A window of lines: each line's nq frequencies with the reference's abscissa
        q = 1 - lambda/lambda0 (SRC/atom_mod.f90:2083) uniform in +-qmax, its rate
        weights wnu = trapezoid(nu)/nu and profile weights wq = trapezoid(q)
        (:2086-2094)."""
        cfg, atom = self.cfg, self.atom
        q = np.linspace(-cfg.qmax, cfg.qmax, cfg.nq)
        q32 = q.astype(np.float32)
        wq = _trapezoid(q32.astype(np.float64))
        nline, nf = len(lines), len(lines) * cfg.nq
        nus = [atom.nu0[l] / (1.0 - q) for l in lines]
        nu = np.concatenate(nus)
        qa = np.zeros((nline, nf), np.float32)
        wqa = np.zeros((nline, nf))
        wnu = np.zeros((nline, nf))
        iblue = np.zeros(nline, np.int32)
        ired = np.zeros(nline, np.int32)
        for k, l in enumerate(lines):
            fs = k * cfg.nq
            iblue[k], ired[k] = fs + 1, fs + cfg.nq
            qa[k, fs:fs + cfg.nq] = q32
            wqa[k, fs:fs + cfg.nq] = wq
            wnu[k, fs:fs + cfg.nq] = _trapezoid(nus[k]) / nus[k]
        li = atom.ti[lines].astype(np.int32)
        lj = atom.tj[lines].astype(np.int32)
        gi, gj = atom.g[li - 1], atom.g[lj - 1]
        Aij = atom.Aji[lines] * gj / gi
        # opacity per unit profile in (nu - nu0)/nu0: chi_line = n_i Bij phi (b_i - b_j beta)
        Bij = PI_E2_MC * atom.f[lines] / atom.nu0[lines]
        chi_c, beta, B, gB = self._background(nu)
        d = lambda a, k: torch.as_tensor(np.ascontiguousarray(a), dtype=m3drt.BB_DTYPES[k], device=self.device)
        const = dict(iblue=d(iblue, "iblue"), ired=d(ired, "ired"), lev_i=d(li, "lev_i"), lev_j=d(lj, "lev_j"),
                     Aij=d(Aij, "Aij"), Bij=d(Bij, "Bij"), lower_off=d(np.zeros(nline), "lower_off"),
                     q=d(qa, "q"), wnu=d(wnu, "wnu"), wq=d(wqa, "wq"), g=d(atom.g, "g"))
        return _Window(lines, nf, chi_c, beta, B, gB, const)

    # ---------------------------------------------------------------- sweeps
    def _sweep(self, r, chi, S, ghost_B, **out):
        """Direction r over a window's (nf, ncell) chi, S (sweep dtype)."""
        cfg, box = self.cfg, self.box
        nf = chi.shape[0]
        v = self.vec[r]
        up = v[2] > 0
        shp = (nf, cfg.nz, cfg.ny, cfg.nx)
        if up:
            # set_upstream_boundary / lte_ghost (SRC/rt_patch_mod.f90:3352-3373): the LTE
            # plane below, chi = chi(z=1) chi500_ghost / chi500(1), S = I = B(T_ghost)
            gchi = (chi[:, :cfg.nx * cfg.ny].double() * self.r500[None]).to(chi.dtype).contiguous()
            gS = ghost_B.to(chi.dtype)
            gI = gS
        else:
            gchi = torch.zeros((nf, cfg.nx * cfg.ny), dtype=chi.dtype, device=self.device)
            gS = gI = gchi
        pl = lambda t: t.view(nf, cfg.ny, cfg.nx)
        self.ctx.sc_sweep(chi.view(shp), S.view(shp), pl(gchi), pl(gS), pl(gI), box.chi500, self.cu[r],
                          box.g500_bot if up else box.g500_top, dir=1 if up else -1, ds=box.ds, vec=v,
                          layout="fmajor", wj=float(self.wj[r]),
                          **{k: (t.view(shp) if isinstance(t, torch.Tensor) and t.dim() == 2 else t)
                             for k, t in out.items()})

    def _continuum(self):
        """Frequencies, cross sections and the angle-averaged J, PsiJ of the continua."""
        cfg, atom = self.cfg, self.atom
        nl, nc = atom.nline, len(atom.ti) - atom.nline
        nus = [atom.nu0[nl + j] * np.exp(np.linspace(0.0, math.log(3.0), cfg.ncont_f)) for j in range(nc)]
        nu = np.concatenate(nus)
        nf = len(nu)
        al = np.zeros((nc, nf), np.float32)
        wnu = np.zeros((nc, nf))
        for j in range(nc):
            fs = j * cfg.ncont_f
            al[j, fs:fs + cfg.ncont_f] = atom.sigma0[j] * (atom.nu0[nl + j] / nus[j]) ** 3
            # 4 pi / (h nu) dnu (SRC/atom_mod.f90:2053-2055)
            wnu[j, fs:fs + cfg.ncont_f] = _trapezoid(nus[j]) / (H_4PI * nus[j])
        chi_c, beta, B, gB = self._background(nu)
        self.cont = dict(chi_c=chi_c, B=B, ghost_B=gB)
        dt = self.sweep_dtype
        chi, S = chi_c.to(dt), B.to(dt)
        J = torch.zeros_like(chi)
        PsiJ = torch.zeros_like(chi)
        for r in range(len(self.vec)):
            self._sweep(r, chi, S, gB, J=J, PsiJ=PsiJ, get_rates=True)
        d = lambda a, k: torch.as_tensor(np.ascontiguousarray(a), dtype=m3drt.BF_DTYPES[k], device=self.device)
        self.bf = dict(iblue=d(np.arange(nc) * cfg.ncont_f + 1, "iblue"), ired=d((np.arange(nc) + 1) * cfg.ncont_f, "ired"),
                       lev_i=d(atom.ti[nl:], "lev_i"), lev_j=d(atom.tj[nl:], "lev_j"),
                       active=d(np.ones(nc), "active"), ev=d(atom.ev, "ev"), wnu=d(wnu, "wnu"), alpha=d(al, "alpha"),
                       twohnu3_c2=d(TWOH_C2 * nu ** 3, "twohnu3_c2"), e_kT=self.e_kT64, nstar=self.nstar,
                       Jnu=J.float().contiguous(), beta=beta, PsiJ=PsiJ.float().contiguous())

    # ---------------------------------------------------------------- iteration
    def _line_rates(self, w: _Window, gij, gji):
        ctx, nc = self.ctx, self.ncell
        nf = w.nf
        f32 = lambda: torch.empty((nf, nc), dtype=torch.float32, device=self.device)
        i32 = lambda: torch.empty((nf, nc), dtype=torch.int32, device=self.device)
        xhat, ehat = f32(), f32()
        sp = dict(split_opac1=f32(), split_opac2=f32(), split_chi1=f32(), split_chi2=f32(), split_id=i32(),
                  split_chi_id=i32())
        I, Psi, Irem, esc = f32(), f32(), f32(), f32()
        sweep64 = self.sweep_dtype == torch.float64
        l0, l1 = w.lines[0], w.lines[-1] + 1
        cells = dict(b=self.b, nstar=self.nstar, pop_ratio=self.pop_ratio, inv_df=self.inv_df, e_kT=self.e_kT32,
                     adamp=self.adamp[l0:l1])
        for r, v in enumerate(self.vec):
            vproj = (self.vz * (v[2] / CC)).float().contiguous()
            d = dict(w.const, vproj=vproj, **cells)
            if self.cfg.fused:
                # the same steps as below in one kernel (m3drt_window_opacity)
                ctx.window_opacity(d, w.chi_c, w.Bnu, w.beta, xhat, ehat, split=sp, split_offset=0)
            else:
                xhat.zero_()
                torch.mul(w.chi_c, w.Bnu, out=ehat)
                ctx.convert_to_work_units(w.chi_c, xhat, ehat, w.beta)
                # the split tracking starts from everything that is not a line of the
                # atom (the continuum) as the one contribution of index 0
                # (SRC/rt_patch_mod.f90:1467-1488)
                sp["split_opac1"].copy_(ehat)
                sp["split_chi1"].copy_(xhat)
                for k in ("split_opac2", "split_chi2", "split_id", "split_chi_id"):
                    sp[k].zero_()
                ctx.bb_opacity(d, xhat, ehat, wmu=float(self.wj[r]), split=sp, split_offset=0)
                ctx.work_to_transfer(xhat, ehat, w.beta, split_opac1=sp["split_opac1"],
                                     split_opac2=sp["split_opac2"])
            chi, S = xhat, ehat
            if sweep64:
                o64 = [t.double() for t in (I, Psi, Irem, esc)]
                self._sweep(r, chi.double(), S.double(), w.ghost_B, I_out=o64[0], Psi_out=o64[1], Irem_out=o64[2],
                            escape_out=o64[3], split_rates=True, get_rates=True)
                for a, b64 in zip((I, Psi, Irem, esc), o64):
                    a.copy_(b64)
            else:
                self._sweep(r, chi, S, w.ghost_B, I_out=I, Psi_out=Psi, Irem_out=Irem, escape_out=esc,
                            split_rates=True, get_rates=True)
            ctx.bb_rates(dict(d, Inu=I, Psi=Psi, Bnu=w.Bnu, beta=w.beta), float(self.wj[r]), 1.0,
                         gij[l0:l1], gji[l0:l1], precond=self.cfg.precond, own_norm=True, largest=True,
                         split=dict(sp, Irem=Irem, escape=esc, chi=chi), split_offset=0)

    def rates(self):
        """The radiative rate sums of one iteration: (gamma_ij, gamma_ji), (nrad, ncell)
        float32, normalised (normalize_profiles).  With cfg.graph the whole pass (every
        window's opacity, sweeps and rate sums, the continua, the normalisation) is
        captured once, after an eager pass that creates the tables and scratch it
        uses, and replayed from then on: the same kernels on the same buffers, whose
        contents (departure coefficients) the iteration changes in place."""
        if not (self.cfg.graph and self.device == "cuda"):
            return self._rates()
        if self._graph is None:
            out = self._rates()
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._graph_out = self._rates()
            self._graph = g
            return out
        self._graph.replay()
        return self._graph_out

    def _rates(self):
        atom = self.atom
        nrad = len(atom.ti)
        gij = torch.zeros((nrad, self.ncell), dtype=torch.float32, device=self.device)
        gji = torch.zeros_like(gij)
        for w in self.windows:
            self._line_rates(w, gij, gji)
        self.bf["b"] = self.b
        self.ctx.bf_rates(self.bf, atom.nline, gij, gji, precond=False)
        self.ctx.normalize_rates(self.atom_dev, self.nstar, gij, gji)
        return gij, gji

    def update(self, gij, gji):
        """update_populations of every cell in place; returns the convergence measure."""
        return self.ctx.rate_solve(self.atom_dev, self.e_kT64, self.temp, self.totn, self.nstar, self.b, gij, gji,
                                   self.cup, sor_w=self.cfg.sor_w, conv_pop_floor=self.cfg.conv_pop_floor)

    def iterate(self):
        """One ALI iteration: the rates, the statistical-equilibrium update and, when
        its gate fires, Ng's extrapolation; returns the update's convergence measure
        (with 'ng', the extrapolation's largest relative change, when it was applied)."""
        gij, gji = self.rates()
        self._iter += 1
        self._oldb = (self._oldb + [self.b.clone()])[-3:]
        res = self.update(gij, gji)
        if self.ng is not None:
            ng = self.ng
            ng.promote()
            fired = False
            if self.cfg.ng_step > 0:
                # the fixed schedule, once a Gram matrix has been promoted
                fire = ng_fixed_fires(self._iter, self.cfg.ng_step, self.cfg.ng_start) and ng.s_done > 0
            else:
                fire = ng.should_fire(self._iter, res["maxerr"], self.cfg.conv)
            if fire:
                change = ng.extrapolate(self.b, self._oldb, self.nstar, self.totn)
                if change is not None:
                    ng.iter_fired = self._iter
                    res["ng"] = change
                    fired = True
            if len(self._oldb) == 3 and not fired:
                ng.accumulate(self.b, self._oldb, res["nclamp"])
        return res

    def solve(self, log=None, iterations=None):
        """Iterate to max |dn/n| < conv or max_iter; returns the iteration count.
        iterations: run exactly that many instead (the Ng gate still sees conv)."""
        for it in range(1, (iterations or self.cfg.max_iter) + 1):
            t0 = time.perf_counter()
            res = self.iterate()
            if self.device == "cuda":
                torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            self.history.append((res["maxerr"], dt))
            self.ng_history.append("ng" in res)
            if log:
                log(f"[ali] iteration {it}: max|dn/n| = {res['maxerr']:.3e}  ({dt:.2f} s)")
            if iterations is None and abs(res["maxerr"]) < self.cfg.conv:
                return it
        return iterations or self.cfg.max_iter


def main():
    """Run the NLTE workload to convergence and print one JSON line (iterations, time
    per iteration, the convergence history); with --fp64-sweeps also the same
    iterations with the plain FP64 sweep, and the departure coefficients' largest
    relative difference between the two."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=60)
    ap.add_argument("--nz", type=int, default=120)
    ap.add_argument("--max-iter", type=int, default=60)
    ap.add_argument("--fp64-sweeps", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch the rate pass eagerly every iteration")
    ap.add_argument("--ng-step", type=int, default=0, help="> 0: Ng on the reference's fixed schedule")
    ap.add_argument("--ng-start", type=int, default=8)
    ap.add_argument("--sor-w", type=float, default=1.0, help="&nlte_params sor_w (1: undamped)")
    ap.add_argument("--ng-gate", action="append", default=[], metavar="NAME=VALUE",
                    help="an automatic Ng gate parameter (&nlte_params ng_<NAME>), e.g. min_gap=3")
    ap.add_argument("--conv-pop-floor", type=float, default=0.0,
                    help="&nlte_params conv_pop_floor: levels below this share of the atoms weigh in the "
                         "convergence measure by their share (0: every level fully)")
    ap.add_argument("--lockstep", action="store_true",
                    help="with --fp64-sweeps: step both runs together for --max-iter iterations and report "
                         "their departure coefficients' difference at every iteration")
    args = ap.parse_args()
    gate = dict(kv.split("=", 1) for kv in args.ng_gate)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    ctx = m3drt.Context(0)
    cfg = AliConfig(nx=args.nx, ny=args.nx, nz=args.nz, max_iter=args.max_iter, graph=not args.no_graph,
                    ng_step=args.ng_step, ng_start=args.ng_start, sor_w=args.sor_w, ng_gate=gate,
                    conv_pop_floor=args.conv_pop_floor)
    t0 = time.perf_counter()
    p = NlteProblem(ctx, cfg)
    torch.cuda.synchronize()
    log(f"[ali] set-up {time.perf_counter() - t0:.1f} s")
    if args.lockstep:
        if not args.fp64_sweeps:
            ap.error("--lockstep compares with the FP64-sweep run: give --fp64-sweeps")
        q = NlteProblem(ctx, cfg, sweep_dtype=torch.float64)
        rows = []
        for it in range(1, cfg.max_iter + 1):
            r32, r64 = p.iterate(), q.iterate()
            d = float((p.b / q.b - 1).abs().max())
            rows.append(dict(iteration=it, maxerr_fp32=r32["maxerr"], maxerr_fp64=r64["maxerr"], ng_fp32="ng" in r32,
                             ng_fp64="ng" in r64, b_max_rel_diff=d))
            log(f"[ali] iteration {it}: max|dn/n| {r32['maxerr']:.3e} (FP32 sweeps) {r64['maxerr']:.3e} (FP64); "
                f"b differ by {d:.3e}")
        ctx.close()
        print(json.dumps(dict(workload="3d_nlte_ali", box=[cfg.nx, cfg.ny, cfg.nz], lockstep=rows)), flush=True)
        return
    n = p.solve(log=log)
    its = [dt for _, dt in p.history]
    out = dict(workload="3d_nlte_ali", box=[cfg.nx, cfg.ny, cfg.nz], levels=p.atom.nk, lines=p.atom.nline,
               continua=len(p.atom.ti) - p.atom.nline, frequencies_per_line=cfg.nq, rays=len(p.vec),
               iterations=n, converged=abs(p.history[-1][0]) < cfg.conv, maxerr=[e for e, _ in p.history],
               s_per_iteration=float(np.median(its)),
               line_updates_per_iteration=p.ncell * p.atom.nline * cfg.nq * len(p.vec),
               b_min=float(p.b.min()), b_max=float(p.b.max()), rate_pass_graph=cfg.graph,
               ng=("fixed schedule" if cfg.ng_step > 0 else "automatic") if cfg.ng else "off",
               ng_step=cfg.ng_step, ng_start=cfg.ng_start, sor_w=cfg.sor_w, ng_gate=cfg.ng_gate,
               conv_pop_floor=cfg.conv_pop_floor,
               ng_fired=[i + 1 for i, h in enumerate(p.ng_history) if h])
    out["line_updates_per_s"] = out["line_updates_per_iteration"] / out["s_per_iteration"]
    if args.fp64_sweeps:
        b32 = p.b.clone()
        del p
        torch.cuda.empty_cache()
        q = NlteProblem(ctx, cfg, sweep_dtype=torch.float64)
        n64 = q.solve(log=log)
        out["fp64_sweeps"] = dict(iterations=n64, s_per_iteration=float(np.median([dt for _, dt in q.history])),
                                  b_max_rel_diff=float((b32 / q.b - 1).abs().max()))
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
