"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py import this module; it is the checker, never the
thing measured or shipped.  The functions restate the reference algorithms of
OpenMULTI3D (experiments/OpenMULTI3D/ of Bergemann-Group-development/dispatch);
each C function cites the reference file:line it follows (oracle/m3d_oracle.c).

Cubes use the reference's Fortran order with frequency fastest.  numpy arrays of
shape (nz, ny, nx, nf) in C order have exactly that memory layout, so the
wrappers take and return them in that shape.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class ScParams(C.Structure):
    _fields_ = [("nf", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dir", C.c_int), ("ds", C.c_double * 3), ("vec", C.c_double * 3),
                ("maser_dtau_max", C.c_double), ("maser_Inu_max", C.c_double),
                ("split_rates", C.c_int), ("get_rates", C.c_int), ("nf_stride", C.c_int)]


class RayT(C.Structure):
    _fields_ = [("xkern", C.c_double * 4), ("ykern", C.c_double * 4), ("dsup", C.c_double),
                ("upo1", C.c_int), ("upo2", C.c_int)]


def build():
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        # make rebuilds the library whenever a source is newer than it, so a
        # stale build is never the checker
        if not os.path.exists(_LIB) or shutil.which("make"):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


def _declare(L):
    vp = C.c_void_p
    for n in ("orc_linear_weights_n", "orc_parab_moments_n"):
        getattr(L, n).argtypes = [C.c_int, f32p, f32p, f32p, f32p, f32p]
        getattr(L, n).restype = None
    L.orc_exp_neg_n.argtypes = [C.c_int, f32p, f32p, f32p]
    L.orc_exp_neg_n.restype = None
    L.orc_cc_kernel.argtypes = [C.c_double, C.POINTER(C.c_double)]
    L.orc_set_stencil.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(RayT)]
    L.orc_gauss_quadrature.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, f64p, f64p]
    L.orc_gauss_quadrature.restype = C.c_int
    for n in ("orc_short_char2_f32", "orc_short_char2_f64"):
        getattr(L, n).argtypes = [C.POINTER(ScParams)] + [vp] * 11
        getattr(L, n).restype = None
    for n in ("orc_long_char2_f32", "orc_long_char2_f64"):
        getattr(L, n).argtypes = [C.POINTER(ScParams)] + [vp] * 10
        getattr(L, n).restype = None
    L.orc_short_char_f32.argtypes = [C.POINTER(ScParams)] + [vp] * 6
    L.orc_long_char_quadratic_f32.argtypes = [C.POINTER(ScParams)] + [vp] * 4
    L.orc_integrate_intensity.argtypes = [C.c_int] * 4 + [C.c_float] + [vp] * 4
    L.orc_sweep_windows_f32.argtypes = [C.POINTER(ScParams), C.c_int] + [vp] * 6 + [C.c_float, vp]
    L.orc_sweep_windows_f32.restype = None


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def linear_weights(t):
    t = np.ascontiguousarray(t, np.float32)
    out = [np.empty_like(t) for _ in range(4)]
    lib().orc_linear_weights_n(t.size, t, *out)
    return tuple(out)  # E, W, Psi, Esc


def parab_moments(t):
    t = np.ascontiguousarray(t, np.float32)
    out = [np.empty_like(t) for _ in range(4)]
    lib().orc_parab_moments_n(t.size, t, *out)
    return tuple(out)  # E, U0, U1, q


def exp_neg(t):
    t = np.ascontiguousarray(t, np.float32)
    E, u0 = np.empty_like(t), np.empty_like(t)
    lib().orc_exp_neg_n(t.size, t, E, u0)
    return E, u0


def cc_kernel(s):
    k = (C.c_double * 4)()
    lib().orc_cc_kernel(float(s), k)
    return np.array(k[:])


def set_stencil(ds, vec):
    r = RayT()
    d = (C.c_double * 3)(*ds)
    v = (C.c_double * 3)(*vec)
    lib().orc_set_stencil(d, v, C.byref(r))
    return dict(xkern=np.array(r.xkern[:]), ykern=np.array(r.ykern[:]), dsup=r.dsup,
                upo1=r.upo1, upo2=r.upo2)


def gauss_quadrature(ntheta, nphi, jnu_weights=True, one_dimensional=False):
    n = 2 * ntheta * nphi
    vec = np.zeros(3 * n)
    wts = np.zeros(n)
    m = lib().orc_gauss_quadrature(ntheta, nphi, int(jnu_weights), int(one_dimensional), vec, wts)
    if m < 0:
        raise ValueError("ntheta out of range")
    return vec[:3 * m].reshape(m, 3), wts[:m]


def _params(nf, nx, ny, nz, dir, ds, vec, maser_dtau_max=2.0, maser_Inu_max=1e30,
            split_rates=False, get_rates=True, nf_stride=0):
    return ScParams(nf, nx, ny, nz, dir, (C.c_double * 3)(*ds), (C.c_double * 3)(*vec),
                    maser_dtau_max, maser_Inu_max, int(split_rates), int(get_rates), nf_stride)


def short_char2(chi, S, I, chi500, chi500_up, g500, dir, ds, vec, eta_row=None, eta_mask=None,
                split_rates=False, get_rates=True, maser_dtau_max=2.0, maser_Inu_max=1e30,
                fp64=False):
    """Restated short_char2 on work cubes of shape (nz+2, ny, nx, nf) including the
    ghost planes; I must hold the upstream ghost plane.  Returns (I, Psi, Irem,
    escape) with Psi/Irem/escape of shape (nz, ny, nx, nf)."""
    dt = np.float64 if fp64 else np.float32
    chi, S = np.ascontiguousarray(chi, dt), np.ascontiguousarray(S, dt)
    I = np.array(I, dt, copy=True, order="C")
    nz2, ny, nx, nf = chi.shape
    nz = nz2 - 2
    c5 = np.ascontiguousarray(chi500, dt)
    cu = np.ascontiguousarray(chi500_up, dt)
    g5 = np.ascontiguousarray(g500, dt)
    psi = np.zeros((nz, ny, nx, nf), dt)
    irem = np.zeros((nz, ny, nx, nf), dt) if split_rates else None
    esc = np.zeros((nz, ny, nx, nf), dt) if split_rates else None
    er = None if eta_row is None else np.ascontiguousarray(eta_row, np.int32)
    em = None if eta_mask is None else np.ascontiguousarray(eta_mask, np.uint8)
    p = _params(nf, nx, ny, nz, dir, ds, vec, maser_dtau_max, maser_Inu_max, split_rates, get_rates)
    fn = lib().orc_short_char2_f64 if fp64 else lib().orc_short_char2_f32
    fn(C.byref(p), _p(chi), _p(S), _p(I), _p(c5), _p(cu), _p(g5), _p(er), _p(em),
       _p(psi), _p(irem), _p(esc))
    return I, psi, irem, esc


def long_char2(chi, S, I, chi500, g500, dir, ds, vec, eta_row=None, eta_mask=None,
               split_rates=False, get_rates=True, maser_dtau_max=2.0, maser_Inu_max=1e30,
               fp64=False):
    dt = np.float64 if fp64 else np.float32
    chi, S = np.ascontiguousarray(chi, dt), np.ascontiguousarray(S, dt)
    I = np.array(I, dt, copy=True, order="C")
    nz2, ny, nx, nf = chi.shape
    nz = nz2 - 2
    c5 = np.ascontiguousarray(chi500, dt)
    g5 = np.ascontiguousarray(g500, dt)
    psi = np.zeros((nz, ny, nx, nf), dt)
    irem = np.zeros((nz, ny, nx, nf), dt) if split_rates else None
    esc = np.zeros((nz, ny, nx, nf), dt) if split_rates else None
    er = None if eta_row is None else np.ascontiguousarray(eta_row, np.int32)
    em = None if eta_mask is None else np.ascontiguousarray(eta_mask, np.uint8)
    p = _params(nf, nx, ny, nz, dir, ds, vec, maser_dtau_max, maser_Inu_max, split_rates, get_rates)
    fn = lib().orc_long_char2_f64 if fp64 else lib().orc_long_char2_f32
    fn(C.byref(p), _p(chi), _p(S), _p(I), _p(c5), _p(g5), _p(er), _p(em), _p(psi), _p(irem), _p(esc))
    return I, psi, irem, esc


def short_char(chi, S, I, dir, ds, vec, split_rates=False, get_rates=True,
               maser_dtau_max=2.0, maser_Inu_max=1e30):
    chi, S = np.ascontiguousarray(chi, np.float32), np.ascontiguousarray(S, np.float32)
    I = np.array(I, np.float32, copy=True, order="C")
    nz2, ny, nx, nf = chi.shape
    nz = nz2 - 2
    psi = np.zeros((nz, ny, nx, nf), np.float32)
    irem = np.zeros_like(psi) if split_rates else None
    esc = np.zeros_like(psi) if split_rates else None
    p = _params(nf, nx, ny, nz, dir, ds, vec, maser_dtau_max, maser_Inu_max, split_rates, get_rates)
    lib().orc_short_char_f32(C.byref(p), _p(chi), _p(S), _p(I), _p(psi), _p(irem), _p(esc))
    return I, psi, irem, esc


def long_char_quadratic(chi, S, I, dir, ds, vec, get_rates=True, maser_dtau_max=2.0,
                        maser_Inu_max=1e30):
    chi, S = np.ascontiguousarray(chi, np.float32), np.ascontiguousarray(S, np.float32)
    I = np.array(I, np.float32, copy=True, order="C")
    nz2, ny, nx, nf = chi.shape
    nz = nz2 - 2
    psi = np.zeros((nz, ny, nx, nf), np.float32)
    p = _params(nf, nx, ny, nz, dir, ds, vec, maser_dtau_max, maser_Inu_max, False, get_rates)
    lib().orc_long_char_quadratic_f32(C.byref(p), _p(chi), _p(S), _p(I), _p(psi))
    return I, psi


def sweep_windows(chi, S, I, chi500, chi500_up, g500, dir, ds, vec, wmu2, J, nwin):
    """CPU baseline: short_char2 + J += wmu2*I over nwin frequency windows on the
    OpenMP threads (OMP_NUM_THREADS).  Arrays in the reference layout; I holds the
    ghost plane and receives the intensities; J (nz, ny, nx, nf) is accumulated."""
    nz2, ny, nx, nf = chi.shape
    p = _params(nf, nx, ny, nz2 - 2, dir, ds, vec, split_rates=False, get_rates=False)
    lib().orc_sweep_windows_f32(C.byref(p), int(nwin), _p(chi), _p(S), _p(I), _p(chi500),
                                _p(chi500_up), _p(g500), float(wmu2), _p(J))


# ---------------------------------------------------------------- QLG / opacity

class QlgCoef(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("ih", "bL", "gG", "cZH", "cL2")]


class QlgCell(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("D", "D2", "SQ", "D2h", "D2G", "D2L")]


class Linelist(C.Structure):
    _fields_ = [("nline", C.c_int), ("llam0", C.c_void_p)] + \
               [(n, C.c_void_p) for n in ("lower_ev", "gpref", "grad", "gw1", "gw2", "gstk",
                                          "unsold", "lev_i", "lev_j", "ipart", "used", "strength",
                                          "reach", "reach_in", "active")]


def _qlg_declare():
    L = lib()
    if getattr(L, "_qlg_ok", False):
        return L
    L.orc_qlg_init.restype = None
    L.orc_qlg_tables.argtypes = [f32p, f32p]
    L.orc_qlg_ref_f64.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, f64p]
    L.orc_qlg_bot_sweep.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, f32p]
    L.orc_qlg_bot_cell.argtypes = [C.c_float, C.POINTER(QlgCell)]
    L.orc_qlg_bot_line_sweep.argtypes = [C.POINTER(QlgCell), C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.c_int, f64p, C.c_double, C.c_double,
                                         C.c_double, C.c_float, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_float), f32p]
    L.orc_qlg_bot_profile.argtypes = [C.c_float] * 4
    L.orc_qlg_bot_profile.restype = C.c_float
    L.orc_add_opacity_equi.argtypes = ([C.POINTER(Linelist), C.c_int, f64p, C.c_int] + [C.c_void_p] * 11
                                       + [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                          C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_float])
    L.orc_add_opacity_equi.restype = None
    L.orc_add_opacity.argtypes = L.orc_add_opacity_equi.argtypes
    L.orc_add_opacity.restype = None
    L._qlg_ok = True
    return L


def qlg_tables():
    L = _qlg_declare()
    tg, th = np.zeros(9020, np.float32), np.zeros(9020, np.float32)
    L.orc_qlg_tables(tg, th)
    return tg, th


def qlg_ref_f64(a, v):
    L = _qlg_declare()
    a = np.atleast_1d(np.asarray(a, np.float64))
    v = np.atleast_1d(np.asarray(v, np.float64))
    out = np.zeros(1)
    res = np.empty(np.broadcast(a, v).shape)
    for i, (ai, vi) in enumerate(np.broadcast(a, v)):
        L.orc_qlg_ref_f64(ai, vi, 0.0, 1, out)
        res.flat[i] = out[0]
    return res


def qlg_bot(a, v):
    """K(a, v) of QLG-bot, point by point (qlg_bot_sweep with n = 1)."""
    L = _qlg_declare()
    a = np.asarray(a, np.float32)
    v = np.asarray(v, np.float32)
    out = np.zeros(1, np.float32)
    res = np.empty(np.broadcast(a, v).shape, np.float32)
    for i, (ai, vi) in enumerate(np.broadcast(a, v)):
        L.orc_qlg_bot_sweep(ai, vi, 0.0, 1, out)
        res.flat[i] = out[0]
    return res


def qlg_line_sweep(D, gL, S, ref, lnq, llam, llc, pos, idl, cap):
    L = _qlg_declare()
    w = QlgCell()
    L.orc_qlg_bot_cell(np.float32(D), C.byref(w))
    nf = len(llam)
    K = np.zeros(nf, np.float32)
    fa, fb, hw = C.c_int(), C.c_int(), C.c_float()
    L.orc_qlg_bot_line_sweep(C.byref(w), gL, S, ref, lnq, nf, np.ascontiguousarray(llam, np.float64),
                             llc, pos, idl, cap, C.byref(fa), C.byref(fb), C.byref(hw), K)
    return fa.value, fb.value, hw.value, K


def add_opacity(lines, llam, chi, eta, cont, logref, cells, lnbase, dep=None, do_nlte=False,
                one_m_eps=None, chi_scat=None, tol=1e-4, line_reject_frac=0.0):
    """Restated add_opacity (a grid of any spacing); arguments as add_opacity_equi."""
    return add_opacity_equi(lines, llam, chi, eta, cont, logref, cells, lnbase, dep, do_nlte, one_m_eps,
                            chi_scat, tol, line_reject_frac, _general=True)


def add_opacity_equi(lines, llam, chi, eta, cont, logref, cells, lnbase, dep=None, do_nlte=False,
                     one_m_eps=None, chi_scat=None, tol=1e-4, line_reject_frac=0.0, _general=False):
    """Restated add_opacity_equi.  lines: dict of SoA arrays (llam0, lower_ev, gpref,
    grad, gw1, gw2, gstk, unsold, lev_i, lev_j, optional ipart, optional reach_in (every
    line's reach: the list's has_reach) and active); chi/eta/cont/logref
    (ncell, nf) float32, updated copies returned; cells: dict of per-cell arrays
    (vproj, idf, vb2, ekt, ne, tothi, spe); lnbase (np, ncell); dep (nlev, ncell)."""
    L = _qlg_declare()
    nl = len(lines["llam0"])
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data_as(C.c_void_p)

    used = np.zeros(nl, np.int32)
    strength = np.zeros(nl, np.float32)
    reach = np.zeros(nl, np.float32)
    ll = Linelist(nl, arr(lines["llam0"], np.float64),
                  *[arr(lines[k], np.float32) for k in ("lower_ev", "gpref", "grad", "gw1", "gw2", "gstk")],
                  *[arr(lines[k], np.int32) for k in ("unsold", "lev_i", "lev_j")],
                  arr(lines["ipart"], np.int32) if lines.get("ipart") is not None else None,
                  used.ctypes.data_as(C.c_void_p), strength.ctypes.data_as(C.c_void_p),
                  reach.ctypes.data_as(C.c_void_p),
                  arr(lines["reach_in"], np.float32) if lines.get("reach_in") is not None else None,
                  arr(lines["active"], np.int32) if lines.get("active") is not None else None)
    chi = np.array(chi, np.float32, copy=True, order="C")
    eta = np.array(eta, np.float32, copy=True, order="C")
    ncell, nf = chi.shape
    lnb = np.ascontiguousarray(lnbase, np.float32)
    np_ = lnb.shape[0]
    dep = np.ones((1, ncell), np.float32) if dep is None else np.ascontiguousarray(dep, np.float32)
    csc = None if chi_scat is None else np.array(chi_scat, np.float32, copy=True, order="C")
    fn = L.orc_add_opacity if _general else L.orc_add_opacity_equi
    fn(C.byref(ll), nf, np.ascontiguousarray(llam, np.float64), ncell,
                           chi.ctypes.data_as(C.c_void_p), eta.ctypes.data_as(C.c_void_p),
                           arr(cont, np.float32), arr(logref, np.float32),
                           *[arr(cells[k], np.float32) for k in ("vproj", "idf", "vb2", "ekt", "ne",
                                                                  "tothi", "spe")],
                           np_, lnb.ctypes.data_as(C.c_void_p), dep.shape[0],
                           dep.ctypes.data_as(C.c_void_p), int(do_nlte),
                           None if one_m_eps is None else arr(one_m_eps, np.float32),
                           None if csc is None else csc.ctypes.data_as(C.c_void_p),
                           int(csc is not None), tol, line_reject_frac)
    return dict(chi=chi, eta=eta, used=used, strength=strength, reach=reach, chi_scat=csc)


# ---------------------------------------------------------------- statistical equilibrium

class OrcAtom(C.Structure):
    _fields_ = [("nk", C.c_int), ("nrad", C.c_int), ("nline", C.c_int), ("npair", C.c_int)] + \
               [(n, C.c_void_p) for n in ("ti", "tj", "active", "g", "ev", "Aji", "pi", "pj")]


class SeParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("max_update", "min_pop_frac", "conv_pop_floor", "sor_w",
                                          "cswitch", "cut_x", "bf_cut_rate", "tmax_update")]


class SeResult(C.Structure):
    _fields_ = [("maxerr", C.c_double), ("epos", C.c_int), ("elvl", C.c_int), ("nclamp", C.c_long),
                ("lvlerr", C.c_void_p)]


SE_DEFAULTS = dict(max_update=1e2, min_pop_frac=1e-30, conv_pop_floor=0.0, sor_w=1.0, cswitch=1.0,
                   cut_x=70.0, bf_cut_rate=1.0, tmax_update=1e7)


def _se_declare():
    L = lib()
    if not getattr(L, "_se_ok", False):
        vp = C.c_void_p
        L.orc_lineq_no_pivot.argtypes = [vp, vp, C.c_int]
        L.orc_lineq_no_pivot.restype = None
        L.orc_normalize_rates.argtypes = [C.POINTER(OrcAtom), C.c_int, vp, vp, vp, vp]
        L.orc_normalize_rates.restype = None
        L.orc_transition_matrix.argtypes = [C.POINTER(OrcAtom), C.POINTER(SeParams), C.c_int, C.c_int] + [vp] * 9
        L.orc_transition_matrix.restype = None
        L.orc_update_populations.argtypes = [C.POINTER(OrcAtom), C.POINTER(SeParams), C.c_int] + [vp] * 8 + \
                                             [C.POINTER(SeResult)]
        L.orc_update_populations.restype = C.c_int
        L._se_ok = True
    return L


class _Keep(list):
    def p(self, a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dt)
        self.append(a)
        return a.ctypes.data_as(C.c_void_p)


def _atom(atom, keep, active=None):
    nrad = len(atom.ti)
    act = np.ones(nrad, np.int32) if active is None else active
    return OrcAtom(atom.nk, nrad, atom.nline, len(atom.pi), keep.p(atom.ti, np.int32), keep.p(atom.tj, np.int32),
                   keep.p(act, np.int32), keep.p(atom.g, np.float64), keep.p(atom.ev, np.float64),
                   keep.p(atom.Aji, np.float64), keep.p(atom.pi, np.int32), keep.p(atom.pj, np.int32))


def lineq_no_pivot(A, b):
    L = _se_declare()
    a = np.asfortranarray(A, np.float64).copy(order="F")
    x = np.array(b, np.float64, copy=True)
    L.orc_lineq_no_pivot(a.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), len(x))
    return x


def normalize_rates(atom, nstar, gij, gji, wphi=None):
    L = _se_declare()
    keep = _Keep()
    at = _atom(atom, keep)
    gij = np.array(gij, np.float32, copy=True)
    gji = np.array(gji, np.float32, copy=True)
    ncell = gij.shape[1]
    L.orc_normalize_rates(C.byref(at), ncell, keep.p(nstar, np.float64), keep.p(wphi, np.float32),
                          gij.ctypes.data_as(C.c_void_p), gji.ctypes.data_as(C.c_void_p))
    return gij, gji


def transition_matrix(atom, cell, e_kT, totn, nstar, b, gij, gji, cup=None, **params):
    L = _se_declare()
    keep = _Keep()
    at = _atom(atom, keep)
    prm = SeParams(**{**SE_DEFAULTS, **params})
    nk = atom.nk
    w = np.zeros((nk, nk), np.float64, order="F")
    e = np.zeros(nk)
    ncell = np.asarray(totn).size
    L.orc_transition_matrix(C.byref(at), C.byref(prm), ncell, cell, keep.p(e_kT, np.float64),
                            keep.p(totn, np.float64), keep.p(nstar, np.float64), keep.p(b, np.float64),
                            keep.p(gij, np.float32), keep.p(gji, np.float32), keep.p(cup, np.float64),
                            w.ctypes.data_as(C.c_void_p), e.ctypes.data_as(C.c_void_p))
    return w, e


def update_populations(atom, e_kT, temp, totn, nstar, b, gij, gji, cup=None, **params):
    """Restated update_populations over all cells; returns (b, result dict)."""
    L = _se_declare()
    keep = _Keep()
    at = _atom(atom, keep)
    prm = SeParams(**{**SE_DEFAULTS, **params})
    b = np.array(b, np.float64, copy=True)
    lvl = np.zeros(atom.nk)
    res = SeResult(0.0, -1, -1, 0, lvl.ctypes.data_as(C.c_void_p))
    ncell = np.asarray(totn).size
    st = L.orc_update_populations(C.byref(at), C.byref(prm), ncell, keep.p(e_kT, np.float64),
                                  keep.p(temp, np.float64), keep.p(totn, np.float64),
                                  keep.p(nstar, np.float64), b.ctypes.data_as(C.c_void_p),
                                  keep.p(gij, np.float32), keep.p(gji, np.float32),
                                  keep.p(cup, np.float64), C.byref(res))
    return b, dict(status=st, maxerr=res.maxerr, epos=res.epos, elvl=res.elvl, nclamp=res.nclamp,
                   lvlerr=lvl)


# ------------------------------------------------------------------ line rates

class BbArgs(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("nline", "nf", "ncell", "nk", "precond", "own_norm", "largest", "kahan")]
                + [(n, C.c_void_p) for n in ("iblue", "ired", "lev_i", "lev_j", "Aij", "Bij", "lower_off", "q",
                                              "wnu", "wq")]
                + [("wmu", C.c_float), ("wtot", C.c_float)]
                + [(n, C.c_void_p) for n in ("g", "b", "nstar", "pop_ratio", "inv_df", "vproj", "e_kT", "adamp",
                                              "Inu", "Psi", "Bnu", "beta")]
                + [("sf", C.c_size_t), ("sc", C.c_size_t)]
                + [(n, C.c_int) for n in ("split_offset", "use_flags", "do_scat", "npw")]
                + [(n, C.c_void_p) for n in ("Irem", "escape", "chi", "split_opac1", "split_opac2", "split_chi1",
                                              "split_chi2", "split_id", "split_chi_id", "chi_scat", "dJnu",
                                              "ntrans", "psi_flags_in", "psi_flags_out")])


SPLIT_FIELDS = dict(Irem=np.float32, escape=np.float32, chi=np.float32, split_opac1=np.float32,
                    split_opac2=np.float32, split_chi1=np.float32, split_chi2=np.float32, split_id=np.int32,
                    split_chi_id=np.int32, chi_scat=np.float32, dJnu=np.float32, ntrans=np.int32,
                    psi_flags_in=np.int32)

RATE_FIELDS = dict(  # name -> dtype of the line-rate inputs (one direction)
    iblue=np.int32, ired=np.int32, lev_i=np.int32, lev_j=np.int32, Aij=np.float64, Bij=np.float64,
    lower_off=np.float32, q=np.float32, wnu=np.float64, wq=np.float64, g=np.float64, b=np.float64,
    nstar=np.float64, pop_ratio=np.float64, inv_df=np.float32, vproj=np.float32, e_kT=np.float32,
    adamp=np.float32, Inu=np.float32, Psi=np.float32, Bnu=np.float32, beta=np.float32)


def bb_rates(d, wmu, wtot, gamma_ij, gamma_ji, precond=False, own_norm=False, largest=False, kahan=False,
             split=None, split_offset=0, psi_flags=None):
    """bound_bound_rates for one direction (SRC/rt_patch_mod.f90:2852-3142).  d holds
    the arrays of RATE_FIELDS; spectra (Inu, Psi, Bnu, beta) are (nf, ncell).  With
    largest and precond, split holds the SPLIT_FIELDS arrays (the solver's Irem and
    escape, total chi, the tracked contributions; chi_scat and dJnu for line
    scattering; ntrans and psi_flags_in (npw, ncell) for the psi flags).
    Returns the accumulated (gamma_ij, gamma_ji), each (nline, ncell) float32, and
    with psi flags the raised words (npw, ncell)."""
    L = lib()
    if not getattr(L, "_bb_ok", False):
        L.orc_bb_rates.argtypes = [C.POINTER(BbArgs), C.c_void_p, C.c_void_p]
        L.orc_bb_rates.restype = None
        L._bb_ok = True
    keep = _Keep()
    nline, nf = d["q"].shape
    nk, ncell = d["b"].shape
    sp = split or {}
    use_flags = "psi_flags_in" in sp
    npw = sp["psi_flags_in"].shape[0] if use_flags else 0
    flags_out = np.zeros((max(npw, 1), ncell), np.int32)
    a = BbArgs(nline, nf, ncell, nk, int(precond), int(own_norm), int(largest), int(kahan),
               *[keep.p(d.get(k), RATE_FIELDS[k]) for k in ("iblue", "ired", "lev_i", "lev_j", "Aij", "Bij",
                                                           "lower_off", "q", "wnu", "wq")],
               wmu, wtot,
               *[keep.p(d.get(k), RATE_FIELDS[k]) for k in ("g", "b", "nstar", "pop_ratio", "inv_df", "vproj",
                                                           "e_kT", "adamp", "Inu", "Psi", "Bnu", "beta")],
               ncell, 1, split_offset, int(use_flags), int("chi_scat" in sp), npw,
               *[keep.p(sp.get(k), SPLIT_FIELDS[k]) for k in ("Irem", "escape", "chi", "split_opac1", "split_opac2",
                                                             "split_chi1", "split_chi2", "split_id", "split_chi_id",
                                                             "chi_scat", "dJnu", "ntrans", "psi_flags_in")],
               flags_out.ctypes.data_as(C.c_void_p))
    gij = np.array(gamma_ij, np.float32, copy=True)
    gji = np.array(gamma_ji, np.float32, copy=True)
    L.orc_bb_rates(C.byref(a), gij.ctypes.data_as(C.c_void_p), gji.ctypes.data_as(C.c_void_p))
    if use_flags:
        return gij, gji, flags_out
    return gij, gji


class BbopArgs(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("nline", "nf", "ncell", "nk", "lte", "split", "split_offset")]
                + [(n, C.c_void_p) for n in ("iblue", "ired", "lev_i", "lev_j", "Bij", "lower_off", "q", "wq",
                                              "wphi_line", "b", "nstar", "pop_ratio", "inv_df", "vproj", "e_kT",
                                              "adamp")]
                + [("wmu", C.c_float), ("sf", C.c_size_t), ("sc", C.c_size_t)]
                + [(n, C.c_void_p) for n in ("xhat", "ehat", "split_opac1", "split_opac2", "split_chi1",
                                              "split_chi2", "split_id", "split_chi_id", "wphi")])


def bb_opacity(d, xhat, ehat, wmu=0.0, lte=False, split=None, split_offset=0, wphi_line=None, wphi=None):
    """add_bound_bound_opacities (SRC/rt_patch_mod.f90:2689-2831) over (nf, ncell)
    spectra xhat, ehat; split: dict of the tracked contributions (split_opac1/2,
    split_id, split_chi1/2, split_chi_id), updated like xhat and ehat; wphi (nline,
    ncell) gains wmu * sum(phi wq) for the lines wphi_line marks.  Returns copies."""
    L = lib()
    if not getattr(L, "_bbop_ok", False):
        L.orc_bb_opacity.argtypes = [C.POINTER(BbopArgs)]
        L.orc_bb_opacity.restype = None
        L._bbop_ok = True
    keep = _Keep()
    nline, nf = d["q"].shape
    nk, ncell = d["b"].shape
    out = dict(xhat=np.array(xhat, np.float32, copy=True), ehat=np.array(ehat, np.float32, copy=True))
    if split is not None:
        for k in ("split_opac1", "split_opac2", "split_chi1", "split_chi2"):
            out[k] = np.array(split[k], np.float32, copy=True)
        for k in ("split_id", "split_chi_id"):
            out[k] = np.array(split[k], np.int32, copy=True)
    if wphi is not None:
        out["wphi"] = np.array(wphi, np.float32, copy=True)
    ptr = lambda k: out[k].ctypes.data_as(C.c_void_p) if k in out else None
    a = BbopArgs(nline, nf, ncell, nk, int(lte), int(split is not None), split_offset,
                 *[keep.p(d.get(k), RATE_FIELDS[k]) for k in ("iblue", "ired", "lev_i", "lev_j", "Bij", "lower_off",
                                                             "q", "wq")],
                 keep.p(wphi_line, np.int32),
                 *[keep.p(d.get(k), RATE_FIELDS[k]) for k in ("b", "nstar", "pop_ratio", "inv_df", "vproj", "e_kT",
                                                             "adamp")],
                 wmu, ncell, 1,
                 *[ptr(k) for k in ("xhat", "ehat", "split_opac1", "split_opac2", "split_chi1", "split_chi2",
                                    "split_id", "split_chi_id", "wphi")])
    L.orc_bb_opacity(C.byref(a))
    return out


class BfArgs(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("ncont", "nline", "nf", "ncell", "nk", "precond")]
                + [("cut_x", C.c_double)]
                + [(n, C.c_void_p) for n in ("iblue", "ired", "lev_i", "lev_j", "active", "ev", "wnu", "alpha",
                                              "twohnu3_c2", "e_kT", "nstar", "b", "Jnu", "beta", "PsiJ")]
                + [("sf", C.c_size_t), ("sc", C.c_size_t)])


BF_FIELDS = dict(iblue=np.int32, ired=np.int32, lev_i=np.int32, lev_j=np.int32, active=np.int32, ev=np.float64,
                 wnu=np.float64, alpha=np.float32, twohnu3_c2=np.float32, e_kT=np.float64, nstar=np.float64,
                 b=np.float64, Jnu=np.float32, beta=np.float32, PsiJ=np.float32)


def bf_rates(d, nline, gamma_ij, gamma_ji, precond=False, cut_x=70.0):
    """bound_free_rates (SRC/rt_patch_mod.f90:3193-3259): d holds BF_FIELDS, spectra
    (nf, ncell); gamma arrays (nline + ncont, ncell), the continua from row nline."""
    L = lib()
    if not getattr(L, "_bf_ok", False):
        L.orc_bf_rates.argtypes = [C.POINTER(BfArgs), C.c_void_p, C.c_void_p]
        L.orc_bf_rates.restype = None
        L._bf_ok = True
    keep = _Keep()
    ncont, nf = d["alpha"].shape
    nk, ncell = d["nstar"].shape
    a = BfArgs(ncont, nline, nf, ncell, nk, int(precond), cut_x,
               *[keep.p(d.get(k), BF_FIELDS[k]) for k in ("iblue", "ired", "lev_i", "lev_j", "active", "ev", "wnu",
                                                         "alpha", "twohnu3_c2", "e_kT", "nstar", "b", "Jnu", "beta",
                                                         "PsiJ")],
               ncell, 1)
    gij = np.array(gamma_ij, np.float32, copy=True)
    gji = np.array(gamma_ji, np.float32, copy=True)
    L.orc_bf_rates(C.byref(a), gij.ctypes.data_as(C.c_void_p), gji.ctypes.data_as(C.c_void_p))
    return gij, gji


def work_to_transfer(xhat, ehat, beta, split_opac1=None, split_opac2=None, chi_scat=None, dJnu=None,
                     suppress_masers=False, maser_chi_floor=1e-4):
    """work_to_transfer_units (SRC/rt_patch_mod.f90:3571-3651) on flat (f, cell)
    arrays; returns (chi, S, n_floored)."""
    L = lib()
    if not getattr(L, "_units_ok", False):
        vp = C.c_void_p
        L.orc_work_to_transfer.argtypes = [C.c_size_t, vp, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                           C.c_double]
        L.orc_work_to_transfer.restype = C.c_long
        L.orc_convert_to_work_units.argtypes = [C.c_size_t, vp, vp, vp, vp]
        L.orc_convert_to_work_units.restype = None
        L._units_ok = True
    keep = _Keep()
    w1 = np.array(xhat, np.float32, copy=True)
    w2 = np.array(ehat, np.float32, copy=True)
    split = split_opac1 is not None
    scat = chi_scat is not None
    nfl = L.orc_work_to_transfer(w1.size, w1.ctypes.data_as(C.c_void_p), w2.ctypes.data_as(C.c_void_p),
                                 keep.p(beta, np.float32), keep.p(split_opac1, np.float32),
                                 keep.p(split_opac2, np.float32), keep.p(chi_scat, np.float32),
                                 keep.p(dJnu, np.float32), int(split), int(scat), int(suppress_masers),
                                 maser_chi_floor)
    return w1, w2, int(nfl)


def convert_to_work_units(chi_cont, ehat, beta):
    """convert_to_work_units (SRC/rt_patch_mod.f90:3518-3536): returns (xhat, ehat)."""
    work_to_transfer(np.zeros(1, np.float32), np.zeros(1, np.float32), np.ones(1, np.float32))  # declares
    L = lib()
    keep = _Keep()
    x = np.zeros(np.shape(chi_cont), np.float32)
    e = np.array(ehat, np.float32, copy=True)
    L.orc_convert_to_work_units(x.size, keep.p(chi_cont, np.float32), x.ctypes.data_as(C.c_void_p),
                                e.ctypes.data_as(C.c_void_p), keep.p(beta, np.float32))
    return x, e


class EtaArgs(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nf", "nx", "ny", "nz", "dir", "rx", "ry", "ghost")] + \
               [(n, C.c_void_p) for n in ("chi", "S", "beta", "gchi", "gS", "src", "mask", "gsrc")]


def update_eta_mask(chi, S, beta, gchi, gS, src, mask, gsrc, dir, rx, ry, ghost=True):
    """update_eta_mask (SRC/rt_patch_mod.f90:3671-3802) on dense flags: cubes
    (nz, nf, ny, nx), ghost plane (nf, ny, nx); returns updated (src, mask, gsrc)."""
    L = lib()
    if not getattr(L, "_eta_ok", False):
        L.orc_update_eta_mask.argtypes = [C.POINTER(EtaArgs)]
        L.orc_update_eta_mask.restype = None
        L._eta_ok = True
    keep = _Keep()
    nz, nf, ny, nx = np.shape(chi)
    src = np.array(src, np.uint8, copy=True)
    mask = np.array(mask, np.uint8, copy=True)
    gsrc = np.array(gsrc, np.uint8, copy=True)
    a = EtaArgs(nf, nx, ny, nz, dir, rx, ry, int(ghost), keep.p(chi, np.float32), keep.p(S, np.float32),
                keep.p(beta, np.float32), keep.p(gchi, np.float32), keep.p(gS, np.float32),
                src.ctypes.data_as(C.c_void_p), mask.ctypes.data_as(C.c_void_p), gsrc.ctypes.data_as(C.c_void_p))
    L.orc_update_eta_mask(C.byref(a))
    return src, mask, gsrc


def background_emission(J, B, chi_cont, sigma, PsiJ=None, Jali_old=None, j_to_b=False, wscat=1.0,
                        scat_op_max=0.9):
    """lte_background_emission (SRC/rt_patch_mod.f90:2127-2328) on flat arrays.
    Returns dict(J, Jali, eta, dJnu, n_neg, op_raw, op_max, amp_max)."""
    L = lib()
    if not getattr(L, "_bge_ok", False):
        vp = C.c_void_p
        L.orc_background_emission.argtypes = [C.c_size_t] + [vp] * 9 + [C.c_int, C.c_int, C.c_float, C.c_double, vp]
        L.orc_background_emission.restype = C.c_long
        L._bge_ok = True
    keep = _Keep()
    Jo = np.array(J, np.float32, copy=True)
    n = Jo.size
    ali = PsiJ is not None
    Jali = np.zeros(n, np.float32)
    eta = np.zeros(n, np.float32)
    dJ = np.zeros(n, np.float32)
    diag = np.zeros(3)
    nneg = L.orc_background_emission(n, Jo.ctypes.data_as(C.c_void_p), keep.p(B, np.float32),
                                     keep.p(chi_cont, np.float32), keep.p(sigma, np.float32),
                                     keep.p(PsiJ, np.float32), keep.p(Jali_old, np.float32),
                                     Jali.ctypes.data_as(C.c_void_p), eta.ctypes.data_as(C.c_void_p),
                                     dJ.ctypes.data_as(C.c_void_p), int(ali), int(j_to_b), wscat, scat_op_max,
                                     diag.ctypes.data_as(C.c_void_p))
    return dict(J=Jo, Jali=Jali, eta=eta, dJnu=dJ, n_neg=int(nneg), op_raw=diag[0], op_max=diag[1], amp_max=diag[2])


def short_char_hermite(chi, S, I, beta, pe4, temp, dpe, dlnt, hknu, dir, ds, vec, split_rates=False,
                       get_rates=True, maser_dtau_max=2.0, maser_Inu_max=1e30):
    """short_char_hermite (SRC/rt_solvers_mod.f90:1917-2248): work arrays (nz+2, ny, nx,
    nf) with ghosts; beta (nz, ny, nx, nf); pe4, temp (nz, ny, nx); dpe, dlnt tuples
    of three (nz, ny, nx) gradients; hknu (nf).  Returns (I, Psi, Irem, escape)."""
    L = lib()
    if not getattr(L, "_herm_ok", False):
        L.orc_short_char_hermite_f32.argtypes = [C.POINTER(ScParams)] + [C.c_void_p] * 16
        L.orc_short_char_hermite_f32.restype = None
        L._herm_ok = True
    keep = _Keep()
    chi, S = np.ascontiguousarray(chi, np.float32), np.ascontiguousarray(S, np.float32)
    I = np.array(I, np.float32, copy=True, order="C")
    nz2, ny, nx, nf = chi.shape
    nz = nz2 - 2
    psi = np.zeros((nz, ny, nx, nf), np.float32)
    irem = np.zeros_like(psi) if split_rates else None
    esc = np.zeros_like(psi) if split_rates else None
    p = _params(nf, nx, ny, nz, dir, ds, vec, maser_dtau_max, maser_Inu_max, split_rates, get_rates)
    f32 = lambda a: keep.p(a, np.float32)
    L.orc_short_char_hermite_f32(C.byref(p), _p(chi), _p(S), _p(I), f32(beta), f32(pe4), f32(temp),
                                 f32(dpe[0]), f32(dpe[1]), f32(dpe[2]), f32(dlnt[0]), f32(dlnt[1]), f32(dlnt[2]),
                                 f32(hknu), _p(psi), _p(irem), _p(esc))
    return I, psi, irem, esc


LCH_GHOST = {"top": 0, "bottom": 1, "upstream": 2}


def bezier3_weights(t):
    """bezier3_weights (SRC/rt_solvers_mod.f90:1301-1323): (E, w0, w1, w2, w3)."""
    L = lib()
    if not getattr(L, "_b3_ok", False):
        L.orc_bezier3_weights_n.argtypes = [C.c_int] + [C.c_void_p] * 6
        L.orc_bezier3_weights_n.restype = None
        L._b3_ok = True
    t = np.ascontiguousarray(t, np.float32)
    out = [np.empty_like(t) for _ in range(5)]
    L.orc_bezier3_weights_n(t.size, _p(t), *[_p(o) for o in out])
    return tuple(out)


def long_char_hermite(chi, S, I, beta, pe4, temp, dpe_dz, dlnt_dz, hknu, dir, ds, vec, ghost="top",
                      gpe=None, gdpe=None, gtemp=None, gdlnt=None, get_rates=True, maser_dtau_max=2.0,
                      maser_Inu_max=1e30):
    """long_char_hermite (SRC/rt_solvers_mod.f90:1052-1504) along every column: work
    arrays (nz+2, ny, nx, nf) with ghosts; beta (nz, ny, nx, nf); pe4, temp, dpe_dz,
    dlnt_dz (nz, ny, nx), gradients per cm along +z; hknu (nf); ghost "top" (zeroed
    top ghost), "bottom" (LTE ghost, the model's Pe/T below) or "upstream" (a
    neighbour's last plane) with its thermodynamics gpe, gdpe, gtemp, gdlnt (ny, nx).
    Returns (I, Psi)."""
    L = lib()
    if not getattr(L, "_lch_ok", False):
        L.orc_long_char_hermite_f32.argtypes = [C.POINTER(ScParams)] + [C.c_void_p] * 9 + [C.c_int] + \
            [C.c_void_p] * 5
        L.orc_long_char_hermite_f32.restype = None
        L._lch_ok = True
    keep = _Keep()
    chi, S = np.ascontiguousarray(chi, np.float32), np.ascontiguousarray(S, np.float32)
    I = np.array(I, np.float32, copy=True, order="C")
    nz2, ny, nx, nf = chi.shape
    nz = nz2 - 2
    if nz < 2:
        raise ValueError("long_char_hermite needs nz >= 2")
    kind = LCH_GHOST[ghost]
    if kind and any(a is None for a in (gpe, gdpe, gtemp, gdlnt)):
        raise ValueError("the bottom and upstream ghosts need gpe, gdpe, gtemp and gdlnt")
    psi = np.zeros((nz, ny, nx, nf), np.float32)
    p = _params(nf, nx, ny, nz, dir, ds, vec, maser_dtau_max, maser_Inu_max, False, get_rates)
    f32 = lambda a: keep.p(a, np.float32)
    L.orc_long_char_hermite_f32(C.byref(p), _p(chi), _p(S), _p(I), f32(beta), f32(pe4), f32(temp), f32(dpe_dz),
                                f32(dlnt_dz), f32(hknu), kind, f32(gpe), f32(gdpe), f32(gtemp), f32(gdlnt),
                                _p(psi))
    return I, psi


def long_char_linear(chi, S, I, dir, ds, vec, split_rates=False, get_rates=True, maser_dtau_max=2.0,
                     maser_Inu_max=1e30):
    """long_char_linear (SRC/rt_solvers_mod.f90:666-757) along every column: work
    arrays (nz+2, ny, nx, nf) with ghosts.  Returns (I, Psi, Irem, escape)."""
    L = lib()
    if not getattr(L, "_lcl_ok", False):
        L.orc_long_char_linear_f32.argtypes = [C.POINTER(ScParams)] + [C.c_void_p] * 6
        L.orc_long_char_linear_f32.restype = None
        L._lcl_ok = True
    chi, S = np.ascontiguousarray(chi, np.float32), np.ascontiguousarray(S, np.float32)
    I = np.array(I, np.float32, copy=True, order="C")
    nz2, ny, nx, nf = chi.shape
    nz = nz2 - 2
    psi = np.zeros((nz, ny, nx, nf), np.float32)
    irem = np.zeros_like(psi) if split_rates else None
    esc = np.zeros_like(psi) if split_rates else None
    p = _params(nf, nx, ny, nz, dir, ds, vec, maser_dtau_max, maser_Inu_max, split_rates, get_rates)
    L.orc_long_char_linear_f32(C.byref(p), _p(chi), _p(S), _p(I), _p(psi), _p(irem), _p(esc))
    return I, psi, irem, esc


def long_char_hermite2(chi, S, I, pe4, dpe_dz, dir, ds, vec, ghost="top", gpe=None, gdpe=None, get_rates=True,
                       maser_dtau_max=2.0, maser_Inu_max=1e30):
    """long_char_hermite2 (SRC/rt_solvers_mod.f90:1536-1895) along every column: work
    arrays (nz+2, ny, nx, nf); pe4, dpe_dz (nz, ny, nx); the ghost's gpe, gdpe (ny, nx)
    for "bottom" / "upstream".  Returns (I, Psi)."""
    L = lib()
    if not getattr(L, "_lch2_ok", False):
        L.orc_long_char_hermite2_f32.argtypes = [C.POINTER(ScParams)] + [C.c_void_p] * 5 + [C.c_int] + \
            [C.c_void_p] * 3
        L.orc_long_char_hermite2_f32.restype = None
        L._lch2_ok = True
    keep = _Keep()
    chi, S = np.ascontiguousarray(chi, np.float32), np.ascontiguousarray(S, np.float32)
    I = np.array(I, np.float32, copy=True, order="C")
    nz2, ny, nx, nf = chi.shape
    nz = nz2 - 2
    if nz < 2:
        raise ValueError("long_char_hermite2 needs nz >= 2")
    kind = LCH_GHOST[ghost]
    if kind and (gpe is None or gdpe is None):
        raise ValueError("the bottom and upstream ghosts need gpe and gdpe")
    psi = np.zeros((nz, ny, nx, nf), np.float32)
    p = _params(nf, nx, ny, nz, dir, ds, vec, maser_dtau_max, maser_Inu_max, False, get_rates)
    f32 = lambda a: keep.p(a, np.float32)
    L.orc_long_char_hermite2_f32(C.byref(p), _p(chi), _p(S), _p(I), f32(pe4), f32(dpe_dz), kind, f32(gpe), f32(gdpe),
                                 _p(psi))
    return I, psi


def short_char2_adjoint(chi, S, I, chi500, chi500_up, g500, Ibar, dir, ds, vec, maser_dtau_max=2.0,
                        maser_Inu_max=1e30, chibar=None, Sbar=None):
    """Reverse of short_char2's S step in FP64 (oracle/m3d_adjoint.c): work arrays
    (nz+2, ny, nx, nf) with ghosts; I the forward solution of short_char2(fp64=True);
    Ibar the direct cotangents of I (planes 1..nz).  Returns (Ibar total incl. the
    upstream ghost plane, chibar, Sbar), the last two accumulated onto the given ones."""
    L = lib()
    if not getattr(L, "_adj_ok", False):
        L.orc_short_char2_adjoint_f64.argtypes = [C.POINTER(ScParams)] + [C.c_void_p] * 9
        L.orc_short_char2_adjoint_f64.restype = None
        L.orc_lin_weights_deriv_f64.argtypes = [C.c_double] + [C.POINTER(C.c_double)] * 6
        L.orc_lin_weights_deriv_f64.restype = None
        L._adj_ok = True
    f64 = lambda a: np.ascontiguousarray(a, np.float64)
    chi, S, I = f64(chi), f64(S), f64(I)
    nz2, ny, nx, nf = chi.shape
    Ib = np.array(Ibar, np.float64, copy=True, order="C")
    cb = np.zeros_like(chi) if chibar is None else np.array(chibar, np.float64, copy=True, order="C")
    sb = np.zeros_like(chi) if Sbar is None else np.array(Sbar, np.float64, copy=True, order="C")
    c5, cu, g5 = f64(chi500), f64(chi500_up), f64(g500)
    p = _params(nf, nx, ny, nz2 - 2, dir, ds, vec, maser_dtau_max, maser_Inu_max, False, False)
    L.orc_short_char2_adjoint_f64(C.byref(p), _p(chi), _p(S), _p(I), _p(c5), _p(cu), _p(g5), _p(Ib), _p(cb),
                                  _p(sb))
    return Ib, cb, sb


def lin_weights_deriv(t):
    """(E, W, Psi, dE/dt, dW/dt, dPsi/dt) of the exact linear step at t (FP64)."""
    L = lib()
    L.orc_lin_weights_deriv_f64.argtypes = [C.c_double] + [C.POINTER(C.c_double)] * 6
    L.orc_lin_weights_deriv_f64.restype = None
    out = [C.c_double() for _ in range(6)]
    L.orc_lin_weights_deriv_f64(float(t), *[C.byref(o) for o in out])
    return tuple(o.value for o in out)


def b1d_atmosphere(h, m, lam_ref, want_chi=False):
    """The oracle's synthesis of atmosphere m of a batch given as host numpy arrays
    (the fields of m3drt_b1d_batch): (I_top (nmu, nf), flux (nf)[, chi (nd, nf)]),
    FP64."""
    L = lib()
    if not getattr(L, "_b1d_ok", False):
        L.orc_b1d_atmosphere.argtypes = [C.c_int] * 4 + [C.c_double] * 2 + [f32p] * 10 + [C.c_double, f64p, f64p,
                                                                                          C.c_void_p]
        L.orc_b1d_atmosphere.restype = None
        L._b1d_ok = True
    nd, nf, nmu, nl = (int(h[k][m]) for k in ("nd", "nf", "nmu", "nline"))
    d0, l0, r0, m0 = (int(h[k][m]) for k in ("d_off", "l_off", "ld_off", "mu_off"))
    sl = lambda k, a, n, w=1: np.ascontiguousarray(h[k][w * a:w * (a + n)], np.float32)
    I = np.zeros((nmu, nf))
    F = np.zeros(nf)
    chi = np.zeros((nd, nf)) if want_chi else None
    L.orc_b1d_atmosphere(nd, nf, nmu, nl, float(h["lam0"][m]), float(h["dlam"][m]), sl("dtau", d0, nd),
                         sl("hckT", d0, nd), sl("ivth", d0, nd), sl("cont_a", d0, nd), sl("cont_b", d0, nd),
                         sl("line_off", l0, nl), sl("line_cl", l0, nl), sl("line_rec", r0, nd * nl, 4),
                         sl("mu", m0, nmu), sl("wflux", m0, nmu), float(lam_ref), I, F,
                         None if chi is None else chi.ctypes.data_as(C.c_void_p))
    return (I, F, chi) if want_chi else (I, F)


def rotate_array(x, vec, ds, bot2pivot, ray_offset=(0.0, 0.0)):
    """rotate_array of cubes x (nk, nz, ny, nx, nf) float32 -- the reference's
    work(f, x, y, z, k) -- returned as a new array."""
    L = lib()
    if not getattr(L, "_rot_ok", False):
        L.orc_rotate_array.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                       C.POINTER(C.c_double)] + [C.c_int] * 5 + [f32p, f32p]
        L.orc_rotate_array.restype = None
        L._rot_ok = True
    x = np.ascontiguousarray(x, np.float32)
    nk, nz, ny, nx, nf = x.shape
    out = np.empty_like(x)
    d3 = lambda v: (C.c_double * len(v))(*[float(t) for t in v])
    L.orc_rotate_array(d3(vec), d3(ds), float(bot2pivot), d3(ray_offset), nf, nx, ny, nz, nk, x, out)
    return out


# ------------------------------------------------------------- boundaries, background

def lte_background(nu, temp):
    """lte_background_opacities' (beta, Bnu), each (nf, ncell) float32."""
    L = lib()
    if not getattr(L, "_bg2_ok", False):
        L.orc_lte_background.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_lte_background.restype = None
        L._bg2_ok = True
    nu = np.ascontiguousarray(nu, np.float64)
    temp = np.ascontiguousarray(temp, np.float64)
    beta = np.empty((nu.size, temp.size), np.float32)
    B = np.empty_like(beta)
    L.orc_lte_background(nu.size, temp.size, _p(nu), _p(temp), _p(beta), _p(B))
    return beta, B


def lte_ghost(nu, chi1, g500, chi500_1, T_ghost):
    """The LTE ghost plane: (chi, S, I), each (nf, nxy) float32, from the first
    plane's chi (nf, nxy) and the per-cell g500, chi500 of the first plane and the
    ghost temperature (nxy)."""
    L = lib()
    if not getattr(L, "_lg_ok", False):
        L.orc_lte_ghost.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 8
        L.orc_lte_ghost.restype = None
        L._lg_ok = True
    nu = np.ascontiguousarray(nu, np.float64)
    chi1 = np.ascontiguousarray(chi1, np.float32)
    f = lambda a: np.ascontiguousarray(a, np.float32).ravel()
    g, c1, T = f(g500), f(chi500_1), f(T_ghost)
    out = [np.empty_like(chi1) for _ in range(3)]
    L.orc_lte_ghost(nu.size, g.size, _p(nu), _p(chi1), _p(g), _p(c1), _p(T), *[_p(o) for o in out])
    return tuple(out)


def pchip_ghost(y1, y2, y3, ds3):
    """set_upstream_boundary's PCHIP value one plane below the first three (before
    the clamp at 0)."""
    L = lib()
    if not getattr(L, "_pg_ok", False):
        L.orc_pchip_ghost.argtypes = [C.c_int, C.c_double] + [C.c_void_p] * 4
        L.orc_pchip_ghost.restype = None
        L._pg_ok = True
    ys = [np.ascontiguousarray(y, np.float32).ravel() for y in (y1, y2, y3)]
    out = np.empty_like(ys[0])
    L.orc_pchip_ghost(out.size, float(ds3), *[_p(y) for y in ys], _p(out))
    return out


class OrcOptab(C.Structure):
    _fields_ = [("nnu", C.c_int), ("ntemp", C.c_int), ("nne", C.c_int), ("mu", C.c_double * 2),
                ("rot", C.c_double * 4), ("c0", C.c_double * 2), ("dc", C.c_double * 2), ("lam", C.c_void_p),
                ("val", C.c_void_p), ("scat", C.c_void_p)]


def _optab(lam, val, scat, mu, rot, c0, dc, keep):
    nnu, nt, ne = val.shape
    F = lambda a: keep.p(np.asfortranarray(a, np.float32).ravel(order="F"), np.float32)
    return OrcOptab(nnu, nt, ne, (C.c_double * 2)(*mu), (C.c_double * 4)(*np.asarray(rot, np.float64).ravel(order="F")),
                    (C.c_double * 2)(*c0), (C.c_double * 2)(*dc), keep.p(lam, np.float64), F(val), F(scat))


def optab_slopes(lam, f, mu, rot, c0, dc):
    """sample_build_derivatives' node slopes of f (nnu, ntemp, nne): (dl, dt, de)."""
    L = lib()
    if not getattr(L, "_opt_ok", False):
        L.orc_optab_slopes.argtypes = [C.POINTER(OrcOptab)] + [C.c_void_p] * 4
        L.orc_optab_slopes.restype = None
        L.orc_optab_interpolate.argtypes = [C.POINTER(OrcOptab)] + [C.c_void_p] * 6 + [C.c_int] + [C.c_void_p] * 3 + \
                                           [C.c_int] + [C.c_void_p] * 3
        L.orc_optab_interpolate.restype = None
        L._opt_ok = True
    keep = _Keep()
    t = _optab(lam, f, f, mu, rot, c0, dc, keep)
    fF = np.asfortranarray(f, np.float32).ravel(order="F")
    out = [np.zeros_like(fF) for _ in range(3)]
    L.orc_optab_slopes(C.byref(t), _p(fF), *[_p(o) for o in out])
    return tuple(o.reshape(f.shape, order="F") for o in out)


def optab_interpolate(lam, val, scat, mu, rot, c0, dc, temp, ne, chi500, lam_q):
    """sample_interpolate over cells: (back, scat), each (ncell, nf) float32, and the
    slopes it used (dl, dt, de, sl, st, se)."""
    sv = optab_slopes(lam, val, mu, rot, c0, dc)
    ss = optab_slopes(lam, scat, mu, rot, c0, dc)
    L = lib()
    keep = _Keep()
    t = _optab(lam, val, scat, mu, rot, c0, dc, keep)
    sl = [keep.p(np.asfortranarray(a).ravel(order="F"), np.float32) for a in sv + ss]
    temp = np.ascontiguousarray(temp, np.float64)
    lam_q = np.ascontiguousarray(lam_q, np.float64)
    back = np.zeros((temp.size, lam_q.size), np.float32)
    sc = np.zeros_like(back)
    L.orc_optab_interpolate(C.byref(t), *sl, temp.size, _p(temp), keep.p(ne, np.float64), keep.p(chi500, np.float32),
                            lam_q.size, _p(lam_q), _p(back), _p(sc))
    return back, sc, sv + ss


# ------------------------------------------------------------ line-list composition

def particle_base(temp, vmic, n, part, awgt, frac=1.0, scale=1.0):
    """add_particle_opacity's per-isotope (inv_df, vbar2, lnbase), each (ncell) float32."""
    L = lib()
    if not getattr(L, "_pb_ok", False):
        L.orc_particle_base.argtypes = [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                        C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_particle_base.restype = None
        L.orc_logref.argtypes = [C.c_long, C.c_void_p, C.c_float, C.c_void_p]
        L.orc_logref.restype = None
        L._pb_ok = True
    keep = _Keep()
    nc = np.asarray(temp).size
    out = [np.empty(nc, np.float32) for _ in range(3)]
    L.orc_particle_base(nc, keep.p(temp, np.float32), keep.p(vmic, np.float32), keep.p(n, np.float64), float(scale),
                        keep.p(part, np.float64), float(awgt), float(frac), *[_p(o) for o in out])
    return tuple(out)


def logref_of(chi, tolerance=1e-4):
    lib()
    particle_base(np.ones(1), np.zeros(1), np.ones(1), np.zeros(1), 1.0)   # declares orc_logref
    chi = np.ascontiguousarray(chi, np.float32)
    out = np.empty_like(chi)
    lib().orc_logref(chi.size, _p(chi), np.float32(tolerance), _p(out))
    return out


def add_linelist_opacities(species, llam, all_chi, all_eta, cells, equidistant=True, one_m_eps=None,
                           chi_scat=None, beta=None, tolerance=1e-4):
    """add_linelist_opacities for metals and molecules (HOP/composition_mod.f90:1663-1954):
    species = list of dicts (isotopes: list of (lines dict, awgt, frac); n, part (ncell)
    float64; scale; dep (nlev, ncell) or None; scat_group); all_chi, all_eta, chi_scat,
    beta (ncell, nf); one_m_eps (2, ncell, nf); cells: temp, vmic, vproj, ekt, ne,
    tothi, spe (ncell).  Returns (all_chi, all_eta, chi_scat, chi_l, eta_l)."""
    all_chi = np.array(all_chi, np.float32, copy=True)
    all_eta = np.array(all_eta, np.float32, copy=True)
    logref = logref_of(all_chi, tolerance)
    chi_l = np.zeros_like(all_chi)
    eta_l = np.zeros_like(all_chi)
    csc = None if chi_scat is None else np.array(chi_scat, np.float32, copy=True)
    for sp in species:
        for lines, awgt, frac in sp["isotopes"]:
            if len(lines["llam0"]) == 0:
                continue
            idf, vb2, lnb = particle_base(cells["temp"], cells["vmic"], sp["n"], sp["part"], awgt, frac,
                                          sp.get("scale", 1.0))
            cl = dict(vproj=cells["vproj"], idf=idf, vb2=vb2, ekt=cells["ekt"], ne=cells["ne"],
                      tothi=cells["tothi"], spe=cells["spe"])
            dep = sp.get("dep")
            g = sp.get("scat_group", 0)
            scat = dep is None and g and csc is not None
            r = add_opacity_equi(lines, llam, chi_l, eta_l, all_chi, logref, cl, lnb[None, :], dep=dep,
                                 do_nlte=dep is not None, one_m_eps=one_m_eps[g - 1] if scat else None,
                                 chi_scat=csc if scat else None, tol=tolerance, _general=not equidistant)
            chi_l, eta_l = r["chi"], r["eta"]
            if scat:
                csc = r["chi_scat"]
    all_eta = (all_eta + eta_l).astype(np.float32)
    all_chi = (all_chi + chi_l).astype(np.float32)
    if csc is not None and beta is not None:
        csc = (csc * (np.float32(1) - np.asarray(beta, np.float32))).astype(np.float32)
    return all_chi, all_eta, csc, chi_l, eta_l
