"""Python binding of libm3drt (include/m3drt.h) for PyTorch-ROCm callers.

This is how tumult-side Python reaches the same kernels that dispatch.x reaches
through the Fortran module (dispatch_amd/fortran/m3drt_mod.f90): ctypes over the
C ABI, with torch tensors providing device memory and the current HIP stream.
There is no CPU fallback: a missing library raises, and so does a failing call.

Layouts (element strides are passed to the kernels, every layout works):
  native     tensor shape (nz, nf, ny, nx)   x fastest, then y, f, z  (fast path)
  reference  tensor shape (nz, ny, nx, nf)   work(f,x,y,z), f fastest
  fmajor     tensor shape (nf, nz, ny, nx)   a frequency's cube contiguous: the (f, cell)
                                             spectra of the opacity and rate kernels
Per-cell arrays (chi500 ...) are (nz, ny, nx); planes (ghost, exit) are
(nf, ny, nx) native or (ny, nx, nf) reference.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libm3drt.so")

M3DRT_OK = 0
M3DRT_F32, M3DRT_F64 = 0, 1
M3DRT_LINEAR2, M3DRT_QUADRATIC, M3DRT_LINEAR = 0, 1, 2
M3DRT_LONG_LINEAR = 3
SOLVERS = dict(linear2=M3DRT_LINEAR2, quadratic=M3DRT_QUADRATIC, linear=M3DRT_LINEAR, long_linear=M3DRT_LONG_LINEAR)
M3DRT_SPLIT_RATES, M3DRT_GET_RATES = 1, 2


class M3drtError(RuntimeError):
    pass


class SweepDesc(C.Structure):
    _fields_ = [("nf", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("dir", C.c_int32), ("prec", C.c_int32), ("solver", C.c_int32), ("flags", C.c_int32),
                ("ds", C.c_double * 3), ("vec", C.c_double * 3),
                ("maser_dtau_max", C.c_double), ("maser_Inu_max", C.c_double), ("wj", C.c_double),
                ("cs_f", C.c_int64), ("cs_x", C.c_int64), ("cs_y", C.c_int64), ("cs_z", C.c_int64),
                ("gs_f", C.c_int64), ("gs_x", C.c_int64), ("gs_y", C.c_int64), ("eta_ld", C.c_int64),
                ("eta_nrow", C.c_int64)]


class SweepBufs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "chi", "S", "ghost_chi", "ghost_S", "ghost_I", "chi500", "chi500_up", "g500",
        "eta_row", "eta_mask", "J", "PsiJ", "I_out", "Psi_out", "Irem_out", "escape_out", "exit_I")]


class AdjointBufs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("chi", "S", "ghost_chi", "ghost_S", "ghost_I", "chi500", "chi500_up",
                                          "g500", "I", "Jbar", "Ibar", "exit_bar", "chibar", "Sbar",
                                          "ghost_chibar", "ghost_Sbar", "ghost_Ibar")]


class HermiteFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("beta", "pe4", "temp", "dpe_dx", "dpe_dy", "dpe_dz", "dlnt_dx", "dlnt_dy",
                                          "dlnt_dz", "hknu")]


M3DRT_LCH_GHOST = {"top": 0, "bottom": 1, "upstream": 2}


class LcGhost(C.Structure):
    _fields_ = [("kind", C.c_int32)] + [(n, C.c_void_p) for n in ("pe", "dpe_dz", "temp", "dlnt_dz")]


_lib = None


def lib() -> C.CDLL:
    """The loaded libm3drt.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise M3drtError(f"{LIB_PATH} is missing: build it with `python -m dispatch_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.m3drt_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int]
        L.m3drt_destroy.argtypes = [vp]
        L.m3drt_set_stream.argtypes = [vp, vp]
        L.m3drt_synchronize.argtypes = [vp]
        L.m3drt_last_error.argtypes = [C.c_char_p, C.c_int]
        L.m3drt_sc_sweep.argtypes = [vp, C.POINTER(SweepDesc), C.POINTER(SweepBufs)]
        L.m3drt_native_strides.argtypes = [C.POINTER(SweepDesc)]
        L.m3drt_reference_strides.argtypes = [C.POINTER(SweepDesc)]
        for n in ("m3drt_create", "m3drt_destroy", "m3drt_set_stream", "m3drt_synchronize",
                  "m3drt_last_error", "m3drt_sc_sweep", "m3drt_native_strides",
                  "m3drt_reference_strides"):
            getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().m3drt_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def _check(status: int, what: str):
    if status != M3DRT_OK:
        raise M3drtError(f"{what} failed ({status}): {last_error()}")


def _sweep_desc(chi, layout, cubes, planes, chi500, chi500_up, g500, dir, ds, vec, wj, maser_dtau_max,
                maser_Inu_max):
    """The descriptor of one direction's sweep over the cubes shaped like chi in the
    given layout, after checking every array's device, shape and dtype."""
    if chi.dtype not in (torch.float32, torch.float64):
        raise M3drtError("chi must be float32 or float64")
    nz = chi.shape[0]
    if layout == "native":
        _, nf, ny, nx = chi.shape
    elif layout == "reference":
        _, ny, nx, nf = chi.shape
    elif layout == "fmajor":
        nf, nz, ny, nx = chi.shape
    else:
        raise M3drtError(f"unknown layout {layout!r}")
    cubes = [t for t in cubes if t is not None]
    planes = [t for t in planes if t is not None]
    for t in cubes + planes + [chi500, chi500_up, g500]:
        if t is not None and (not t.is_cuda or not t.is_contiguous()):
            raise M3drtError("every array must be a contiguous device tensor")
    for t in cubes:
        if t.dtype != chi.dtype or tuple(t.shape) != tuple(chi.shape):
            raise M3drtError("cubes must share chi's dtype and shape")
    for t in planes:
        if t.dtype != chi.dtype or t.numel() != nf * nx * ny:
            raise M3drtError("planes must be (nf, ny, nx) / (ny, nx, nf) of chi's dtype")
    for t in (chi500, chi500_up):
        if t is not None and (t.dtype != torch.float32 or t.numel() != nx * ny * nz):
            raise M3drtError("chi500 arrays are float32 of nx*ny*nz cells")
    if g500 is not None and (g500.dtype != torch.float32 or g500.numel() != nx * ny):
        raise M3drtError("g500 is float32 of nx*ny cells")
    d = SweepDesc()
    d.nf, d.nx, d.ny, d.nz = nf, nx, ny, nz
    d.dir = int(dir)
    d.prec = M3DRT_F64 if chi.dtype == torch.float64 else M3DRT_F32
    d.ds = (C.c_double * 3)(*[float(v) for v in ds])
    d.vec = (C.c_double * 3)(*[float(v) for v in vec])
    d.maser_dtau_max, d.maser_Inu_max, d.wj = maser_dtau_max, maser_Inu_max, float(wj)
    d.eta_ld = 0
    if layout == "fmajor":
        # (nf, nz, ny, nx): each frequency's cube contiguous, cells flat as in the
        # (f, cell) spectra of the opacity and rate kernels; planes as native
        _check(lib().m3drt_native_strides(C.byref(d)), "strides")
        d.cs_f, d.cs_z = nz * ny * nx, ny * nx
    else:
        strides = lib().m3drt_native_strides if layout == "native" else lib().m3drt_reference_strides
        _check(strides(C.byref(d)), "strides")
    return d, (nf, nz, ny, nx)


class Context:
    """One m3drt context on a device; calls run on torch's current HIP stream."""

    def __init__(self, device: int = 0):
        self.device = device
        self._h = C.c_void_p()
        _check(lib().m3drt_create(C.byref(self._h), device, 0), "m3drt_create")

    def close(self):
        if self._h:
            lib().m3drt_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bind_stream(self):
        s = torch.cuda.current_stream(self.device)
        _check(lib().m3drt_set_stream(self._h, C.c_void_p(s.cuda_stream)), "m3drt_set_stream")

    def synchronize(self):
        _check(lib().m3drt_synchronize(self._h), "m3drt_synchronize")

    def sc_sweep(self, chi, S, ghost_chi, ghost_S, ghost_I, chi500, chi500_up, g500, *, dir, ds, vec,
                 layout="native", wj=0.0, J=None, PsiJ=None, I_out=None, Psi_out=None, Irem_out=None,
                 escape_out=None, exit_I=None, eta_row=None, eta_mask=None, split_rates=False,
                 get_rates=None, maser_dtau_max=2.0, maser_Inu_max=1e30, solver="linear2", hermite=None,
                 noalloc=False):
        """One directional formal solution (short_char2 in 3D, long_char2 when
        nx = ny = 1; solver="quadratic": long_char_quadratic along every column;
        solver="linear": short_char, the Multi3D-comparison preset; solver="long_linear":
        long_char_linear along every column; hermite=dict of
        the composition fields of m3drt_hermite_fields: short_char_hermite)
        with J += wj*I and PsiJ += wj*Psi fused.  noalloc: through
        m3drt_sc_sweep_noalloc (the library's fixed device scratch, no allocation)."""
        if solver not in SOLVERS:
            raise M3drtError(f"unknown solver {solver!r}")
        if get_rates is None:
            get_rates = PsiJ is not None or Psi_out is not None
        d, (nf, nz, ny, nx) = _sweep_desc(chi, layout, [chi, S, J, PsiJ, I_out, Psi_out, Irem_out, escape_out],
                                          [ghost_chi, ghost_S, ghost_I, exit_I], chi500, chi500_up, g500,
                                          dir, ds, vec, wj, maser_dtau_max, maser_Inu_max)
        d.solver = SOLVERS[solver]
        d.flags = (M3DRT_SPLIT_RATES if split_rates else 0) | (M3DRT_GET_RATES if get_rates else 0)
        if eta_mask is not None:
            if eta_mask.dtype != torch.uint8 or eta_mask.dim() != 2 or eta_mask.shape[1] < nf:
                raise M3drtError("eta_mask is uint8 of shape (nrow, eta_ld >= nf)")
            d.eta_ld = eta_mask.shape[1]
            d.eta_nrow = eta_mask.shape[0]
        if eta_row is not None and (eta_row.dtype != torch.int32 or eta_row.numel() != nx * ny * nz):
            raise M3drtError("eta_row is int32 of nx*ny*nz cells")
        p = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        b = SweepBufs(p(chi), p(S), p(ghost_chi), p(ghost_S), p(ghost_I), p(chi500), p(chi500_up), p(g500),
                      p(eta_row), p(eta_mask), p(J), p(PsiJ), p(I_out), p(Psi_out), p(Irem_out),
                      p(escape_out), p(exit_I))
        self._bind_stream()
        if hermite is not None:
            L = lib()
            if not getattr(L, "_herm_ok", False):
                L.m3drt_sc_sweep_hermite.argtypes = [C.c_void_p, C.POINTER(SweepDesc), C.POINTER(SweepBufs),
                                                     C.POINTER(HermiteFields)]
                L.m3drt_sc_sweep_hermite.restype = C.c_int
                L._herm_ok = True
            for k, t in hermite.items():
                if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
                    raise M3drtError(f"hermite field {k} must be a contiguous float32 device tensor")
            if tuple(hermite["beta"].shape) != tuple(chi.shape):
                raise M3drtError("hermite beta is a cube shaped like chi")
            h = HermiteFields(*[p(hermite[k]) for k, _ in HermiteFields._fields_])
            _check(L.m3drt_sc_sweep_hermite(self._h, C.byref(d), C.byref(b), C.byref(h)), "m3drt_sc_sweep_hermite")
            return
        if noalloc:
            L = lib()
            if not getattr(L, "_noalloc_ok", False):
                L.m3drt_sc_sweep_noalloc.argtypes = [C.c_void_p, C.POINTER(SweepDesc), C.POINTER(SweepBufs)]
                L.m3drt_sc_sweep_noalloc.restype = C.c_int
                L._noalloc_ok = True
            _check(L.m3drt_sc_sweep_noalloc(self._h, C.byref(d), C.byref(b)), "m3drt_sc_sweep_noalloc")
            return
        _check(lib().m3drt_sc_sweep(self._h, C.byref(d), C.byref(b)), "m3drt_sc_sweep")

    def lc_sweep_hermite(self, chi, S, ghost_chi, ghost_S, ghost_I, *, dir, ds, vec, fields, ghost="top",
                         ghost_fields=None, layout="native", wj=0.0, J=None, PsiJ=None, I_out=None, Psi_out=None,
                         exit_I=None, get_rates=None, maser_dtau_max=2.0, maser_Inu_max=1e30, hybrid=False):
        """long_char_hermite along every column (m3drt_lc_sweep_hermite; hybrid=True:
        long_char_hermite2, m3drt_lc_sweep_hermite2, which reads only pe4 and dpe_dz of
        fields and pe, dpe_dz of ghost_fields) with J += wj*I
        and PsiJ += wj*Psi fused.  fields: dict of m3drt_hermite_fields (beta a cube
        like chi; pe4, temp, dpe_dz, dlnt_dz per cell; hknu); ghost "top", "bottom" or
        "upstream", the last two with ghost_fields = dict(pe, dpe_dz, temp, dlnt_dz) of
        (ny, nx) planes."""
        if ghost not in M3DRT_LCH_GHOST:
            raise M3drtError(f"unknown ghost kind {ghost!r}")
        if get_rates is None:
            get_rates = PsiJ is not None or Psi_out is not None
        d, (nf, nz, ny, nx) = _sweep_desc(chi, layout, [chi, S, J, PsiJ, I_out, Psi_out],
                                          [ghost_chi, ghost_S, ghost_I, exit_I], None, None, None,
                                          dir, ds, vec, wj, maser_dtau_max, maser_Inu_max)
        d.solver = SOLVERS["quadratic"]
        d.flags = M3DRT_GET_RATES if get_rates else 0
        tensors = dict(fields)
        if ghost != "top":
            if ghost_fields is None:
                raise M3drtError("the bottom and upstream ghosts need ghost_fields")
            tensors.update({"ghost_" + k: v for k, v in ghost_fields.items()})
        for k, t in tensors.items():
            if t is not None and (t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous()):
                raise M3drtError(f"{k} must be a contiguous float32 device tensor")
        cellk = ("pe4", "dpe_dz") if hybrid else ("pe4", "temp", "dpe_dz", "dlnt_dz")
        ghostk = ("pe", "dpe_dz") if hybrid else ("pe", "dpe_dz", "temp", "dlnt_dz")
        if not hybrid:
            if tuple(fields["beta"].shape) != tuple(chi.shape):
                raise M3drtError("beta is a cube shaped like chi")
            if fields["hknu"].numel() != nf:
                raise M3drtError("hknu holds nf values")
        for k in cellk:
            if fields[k].numel() != nx * ny * nz:
                raise M3drtError(f"{k} holds nx*ny*nz cells")
        if ghost != "top":
            for k in ghostk:
                if ghost_fields[k].numel() != nx * ny:
                    raise M3drtError(f"ghost {k} holds nx*ny cells")
        p = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        b = SweepBufs(p(chi), p(S), p(ghost_chi), p(ghost_S), p(ghost_I), None, None, None, None, None, p(J), p(PsiJ),
                      p(I_out), p(Psi_out), None, None, p(exit_I))
        h = HermiteFields(*[p(fields.get(k)) for k, _ in HermiteFields._fields_])
        gf = ghost_fields or {}
        g = LcGhost(M3DRT_LCH_GHOST[ghost], *[p(gf.get(k)) for k in ("pe", "dpe_dz", "temp", "dlnt_dz")])
        self._bind_stream()
        L = lib()
        if not getattr(L, "_lch_ok", False):
            for fn in (L.m3drt_lc_sweep_hermite, L.m3drt_lc_sweep_hermite2):
                fn.argtypes = [C.c_void_p, C.POINTER(SweepDesc), C.POINTER(SweepBufs), C.POINTER(HermiteFields),
                               C.POINTER(LcGhost)]
                fn.restype = C.c_int
            L._lch_ok = True
        name = "m3drt_lc_sweep_hermite2" if hybrid else "m3drt_lc_sweep_hermite"
        _check(getattr(L, name)(self._h, C.byref(d), C.byref(b), C.byref(h), C.byref(g)), name)

    def sc_adjoint(self, chi, S, ghost_chi, ghost_S, ghost_I, chi500, chi500_up, g500, I, *, dir, ds, vec,
                   chibar, Sbar, layout="native", wj=0.0, Jbar=None, Ibar=None, exit_bar=None,
                   ghost_chibar=None, ghost_Sbar=None, ghost_Ibar=None, maser_dtau_max=2.0, maser_Inu_max=1e30):
        """Reverse of one direction's short_char2 sweep (m3drt_sc_adjoint): from the
        cotangents of J (the forward added wj*I), of the intensity cube and of the
        exit plane, accumulates (+=) those of chi and S and of the upstream ghost
        plane.  I is the forward's intensity cube (its I_out); chi500 is held fixed."""
        d, _ = _sweep_desc(chi, layout, [chi, S, I, Jbar, Ibar, chibar, Sbar],
                           [ghost_chi, ghost_S, ghost_I, exit_bar, ghost_chibar, ghost_Sbar, ghost_Ibar],
                           chi500, chi500_up, g500, dir, ds, vec, wj, maser_dtau_max, maser_Inu_max)
        d.solver = SOLVERS["linear2"]
        d.flags = 0
        p = lambda t: None if t is None else C.c_void_p(t.data_ptr())
        b = AdjointBufs(p(chi), p(S), p(ghost_chi), p(ghost_S), p(ghost_I), p(chi500), p(chi500_up), p(g500),
                        p(I), p(Jbar), p(Ibar), p(exit_bar), p(chibar), p(Sbar), p(ghost_chibar), p(ghost_Sbar),
                        p(ghost_Ibar))
        self._bind_stream()
        L = lib()
        if not getattr(L, "_adj_ok", False):
            L.m3drt_sc_adjoint.argtypes = [C.c_void_p, C.POINTER(SweepDesc), C.POINTER(AdjointBufs)]
            L.m3drt_sc_adjoint.restype = C.c_int
            L._adj_ok = True
        _check(L.m3drt_sc_adjoint(self._h, C.byref(d), C.byref(b)), "m3drt_sc_adjoint")


# ------------------------------------------------------------------ line opacity

M3DRT_OPAC_NLTE, M3DRT_OPAC_SCAT = 1, 2


class Linelist(C.Structure):
    _fields_ = [("nline", C.c_int32)] + [(n, C.c_void_p) for n in (
        "llam0", "lower_ev", "gpref", "grad", "gw1", "gw2", "gstk", "unsold", "lev_i", "lev_j",
        "ipart", "used", "strength", "reach", "reach_in", "active", "sel_work")]


class OpacityArgs(C.Structure):
    _fields_ = [("nf", C.c_int32), ("ncell", C.c_int32), ("npart", C.c_int32), ("nlev", C.c_int32),
                ("flags", C.c_int32), ("tolerance", C.c_float), ("line_reject_frac", C.c_float),
                ("sf", C.c_int64), ("sc", C.c_int64)] + [(n, C.c_void_p) for n in (
        "llam", "vproj", "inv_df", "vbar2", "e_kT", "ne", "tothi", "stark_pe", "lnbase", "dep",
        "chi", "eta", "cont", "logref", "one_m_eps", "chi_scat")]


def _opacity_lib():
    L = lib()
    if not getattr(L, "_opac_ok", False):
        for n in ("m3drt_line_opacity_equi", "m3drt_line_opacity"):
            getattr(L, n).argtypes = [C.c_void_p, C.POINTER(Linelist), C.POINTER(OpacityArgs)]
            getattr(L, n).restype = C.c_int
        L._opac_ok = True
    return L


def _sel_work(lines, nl):
    """The selection scratch a list with reaches or active flags needs (nline + 1
    int32 on the device), else None; reach_in float32 and active int32 of nline."""
    ri, act = lines.get("reach_in"), lines.get("active")
    if ri is None and act is None:
        return None
    if ri is not None and (ri.dtype != torch.float32 or ri.numel() != nl or not ri.is_cuda):
        raise M3drtError("reach_in is a float32 device tensor of nline values")
    if act is not None and (act.dtype != torch.int32 or act.numel() != nl or not act.is_cuda):
        raise M3drtError("active is an int32 device tensor of nline values")
    dev = (ri if ri is not None else act).device
    return torch.empty(nl + 1, dtype=torch.int32, device=dev)


def _line_opacity_equi(self, lines, llam, chi, eta, cont, logref, cells, lnbase, dep=None,
                       nlte=False, one_m_eps=None, chi_scat=None, layout="native", tolerance=1e-4,
                       line_reject_frac=0.0, used=None, strength=None, reach=None, general=False):
    """add_opacity_equi of one isotope's line list on the device.  lines: dict of
    device tensors (llam0 float64; lower_ev, gpref, grad, gw1, gw2, gstk float32;
    unsold, lev_i, lev_j, optional ipart int32; optional reach_in float32 and active
    int32: select_lines with the list's reaches and active flags).  Spectra (chi, eta, cont, logref,
    one_m_eps, chi_scat) are (nf, ncell) in the native layout (cells at unit
    stride) or (ncell, nf) in the reference layout (f at unit stride).  general: the
    grid's spacing may vary (add_opacity, m3drt_line_opacity)."""
    L = _opacity_lib()
    nl = int(lines["llam0"].numel())
    if layout == "native":
        nf, ncell = chi.shape
        sf, sc = ncell, 1
    elif layout == "reference":
        ncell, nf = chi.shape
        sf, sc = 1, nf
    else:
        raise M3drtError(f"unknown layout {layout!r}")
    for t in (chi, eta, cont, logref, one_m_eps, chi_scat):
        if t is not None and (t.dtype != torch.float32 or tuple(t.shape) != tuple(chi.shape)
                              or not t.is_cuda or not t.is_contiguous()):
            raise M3drtError("spectra must be contiguous float32 device tensors shaped like chi")
    if llam.dtype != torch.float64 or llam.numel() != nf:
        raise M3drtError("llam is float64 with nf nodes")
    npart = lnbase.shape[0]
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    sel = _sel_work(lines, nl)
    ll = Linelist(nl, *[p(lines.get(k)) for k in ("llam0", "lower_ev", "gpref", "grad", "gw1", "gw2",
                                                  "gstk", "unsold", "lev_i", "lev_j", "ipart")],
                  p(used), p(strength), p(reach), p(lines.get("reach_in")), p(lines.get("active")), p(sel))
    flags = (M3DRT_OPAC_NLTE if nlte else 0) | (M3DRT_OPAC_SCAT if chi_scat is not None else 0)
    o = OpacityArgs(nf, ncell, npart, 0 if dep is None else dep.shape[0], flags, tolerance,
                    line_reject_frac, sf, sc, p(llam), *[p(cells[k]) for k in (
                        "vproj", "idf", "vb2", "ekt", "ne", "tothi", "spe")],
                    p(lnbase), p(dep), p(chi), p(eta), p(cont), p(logref), p(one_m_eps), p(chi_scat))
    self._bind_stream()
    fn = "m3drt_line_opacity" if general else "m3drt_line_opacity_equi"
    _check(getattr(L, fn)(self._h, C.byref(ll), C.byref(o)), fn)


def _line_opacity(self, *args, **kw):
    """add_opacity of one isotope's line list on a grid of any spacing; arguments as
    line_opacity_equi."""
    return _line_opacity_equi(self, *args, general=True, **kw)


Context.line_opacity_equi = _line_opacity_equi
Context.line_opacity = _line_opacity


# ------------------------------------------------------------------ statistical equilibrium

class RateSolveArgs(C.Structure):
    _fields_ = [("nk", C.c_int32), ("nrad", C.c_int32), ("nline", C.c_int32), ("npair", C.c_int32),
                ("ncell", C.c_int32)] + [(n, C.c_void_p) for n in (
        "trans_i", "trans_j", "trans_active", "g", "ev", "Aji", "pair_i", "pair_j", "e_kT", "temp", "totn",
        "nstar", "b", "gamma_ij", "gamma_ji", "cup")] + [(n, C.c_double) for n in (
        "max_update", "min_pop_frac", "conv_pop_floor", "sor_w", "cswitch", "cut_x", "bf_cut_rate",
        "tmax_update")]


class RateSolveResult(C.Structure):
    _fields_ = [("maxerr", C.c_double), ("epos", C.c_int32), ("elvl", C.c_int32), ("nclamp", C.c_int64),
                ("lvlerr", C.c_void_p)]


SE_DEFAULTS = dict(max_update=1e2, min_pop_frac=1e-30, conv_pop_floor=0.0, sor_w=1.0, cswitch=1.0,
                   cut_x=70.0, bf_cut_rate=1.0, tmax_update=1e7)


class AtomOnDevice:
    """A model atom's constant arrays on the device (see synth.Atom)."""

    def __init__(self, atom, active=None):
        d = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
        self.nk, self.nline = atom.nk, atom.nline
        self.nrad, self.npair = len(atom.ti), len(atom.pi)
        self.ti = d(atom.ti, torch.int32)
        self.tj = d(atom.tj, torch.int32)
        self.active = d(np.ones(self.nrad) if active is None else active, torch.int32)
        self.g = d(atom.g, torch.float64)
        self.ev = d(atom.ev, torch.float64)
        self.Aji = d(atom.Aji, torch.float64)
        self.pi = d(atom.pi, torch.int32)
        self.pj = d(atom.pj, torch.int32)


def _rate_args(atom, e_kT, temp, totn, nstar, b, gamma_ij, gamma_ji, cup, params):
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    ncell = int(totn.numel())
    for t, dt in ((e_kT, torch.float64), (temp, torch.float64), (totn, torch.float64), (nstar, torch.float64),
                  (b, torch.float64), (gamma_ij, torch.float32), (gamma_ji, torch.float32)):
        if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
            raise M3drtError("rate-solve arrays: contiguous device tensors of the documented dtypes")
    prm = {**SE_DEFAULTS, **params}
    return RateSolveArgs(atom.nk, atom.nrad, atom.nline, atom.npair, ncell, p(atom.ti), p(atom.tj),
                         p(atom.active), p(atom.g), p(atom.ev), p(atom.Aji), p(atom.pi), p(atom.pj),
                         p(e_kT), p(temp), p(totn), p(nstar), p(b), p(gamma_ij), p(gamma_ji), p(cup),
                         *[float(prm[k]) for k in ("max_update", "min_pop_frac", "conv_pop_floor", "sor_w",
                                                    "cswitch", "cut_x", "bf_cut_rate", "tmax_update")])


def _se_lib():
    L = lib()
    if not getattr(L, "_se_ok", False):
        L.m3drt_rate_solve_f64.argtypes = [C.c_void_p, C.POINTER(RateSolveArgs), C.POINTER(RateSolveResult)]
        L.m3drt_rate_solve_f64.restype = C.c_int
        L.m3drt_normalize_rates.argtypes = [C.c_void_p, C.POINTER(RateSolveArgs), C.c_void_p]
        L.m3drt_normalize_rates.restype = C.c_int
        L._se_ok = True
    return L


def _rate_solve(self, atom, e_kT, temp, totn, nstar, b, gamma_ij, gamma_ji, cup=None, **params):
    """update_populations on the device: b is updated in place; returns the
    convergence measure (maxerr, epos, elvl, nclamp, lvlerr)."""
    L = _se_lib()
    a = _rate_args(atom, e_kT, temp, totn, nstar, b, gamma_ij, gamma_ji, cup, params)
    lvl = np.zeros(atom.nk)
    res = RateSolveResult(0.0, -1, -1, 0, lvl.ctypes.data_as(C.c_void_p))
    self._bind_stream()
    _check(L.m3drt_rate_solve_f64(self._h, C.byref(a), C.byref(res)), "m3drt_rate_solve_f64")
    return dict(maxerr=res.maxerr, epos=res.epos, elvl=res.elvl, nclamp=res.nclamp, lvlerr=lvl)


def _normalize_rates(self, atom, nstar, gamma_ij, gamma_ji, wphi=None):
    """normalize_profiles' scaling of the rate sums, in place on the device."""
    L = _se_lib()
    totn = torch.zeros(nstar.shape[1], dtype=torch.float64, device="cuda")
    a = _rate_args(atom, totn, totn, totn, nstar, totn, gamma_ij, gamma_ji, None, {})
    a.b = None
    self._bind_stream()
    _check(L.m3drt_normalize_rates(self._h, C.byref(a), None if wphi is None else C.c_void_p(wphi.data_ptr())),
           "m3drt_normalize_rates")


Context.rate_solve = _rate_solve
Context.normalize_rates = _normalize_rates


# ------------------------------------------------------------------ line rates

M3DRT_RATES_PRECOND, M3DRT_RATES_OWN_NORM, M3DRT_RATES_LARGEST = 1, 2, 4
M3DRT_RATES_SCAT, M3DRT_RATES_FLAGS = 8, 16
M3DRT_BBOP_LTE, M3DRT_BBOP_SPLIT = 1, 2

BB_FIELDS = ("iblue", "ired", "lev_i", "lev_j", "Aij", "Bij", "lower_off", "q", "wnu", "wq", "g", "b",
             "nstar", "pop_ratio", "inv_df", "vproj", "e_kT", "adamp", "Inu", "Psi", "Bnu", "beta")
SPLIT_FIELDS = ("Irem", "escape", "chi", "split_opac1", "split_opac2", "split_chi1", "split_chi2", "split_id",
                "split_chi_id", "chi_scat", "dJnu", "ntrans", "psi_flags_in", "psi_flags_out")
BB_DTYPES = dict(iblue=torch.int32, ired=torch.int32, lev_i=torch.int32, lev_j=torch.int32,
                 Aij=torch.float64, Bij=torch.float64, lower_off=torch.float32, q=torch.float32,
                 wnu=torch.float64, wq=torch.float64, g=torch.float64, b=torch.float64, nstar=torch.float64,
                 pop_ratio=torch.float64, inv_df=torch.float32, vproj=torch.float32, e_kT=torch.float32,
                 adamp=torch.float32, Inu=torch.float32, Psi=torch.float32, Bnu=torch.float32,
                 beta=torch.float32, Irem=torch.float32, escape=torch.float32, chi=torch.float32,
                 split_opac1=torch.float32, split_opac2=torch.float32, split_chi1=torch.float32,
                 split_chi2=torch.float32, split_id=torch.int32, split_chi_id=torch.int32,
                 chi_scat=torch.float32, dJnu=torch.float32, ntrans=torch.int32, psi_flags_in=torch.int32,
                 psi_flags_out=torch.int32, xhat=torch.float32, ehat=torch.float32, wphi=torch.float32,
                 wphi_line=torch.int32)


class BbRatesArgs(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in ("nline", "nf", "ncell", "nk", "flags")]
                + [("wmu", C.c_float), ("wtot", C.c_float), ("sf", C.c_int64), ("sc", C.c_int64)]
                + [(n, C.c_void_p) for n in BB_FIELDS + ("gamma_ij", "gamma_ji")]
                + [("split_offset", C.c_int32), ("npw", C.c_int32)]
                + [(n, C.c_void_p) for n in SPLIT_FIELDS])


class BbOpacityArgs(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in ("nline", "nf", "ncell", "nk", "flags", "split_offset")]
                + [("wmu", C.c_float), ("sf", C.c_int64), ("sc", C.c_int64)]
                + [(n, C.c_void_p) for n in ("iblue", "ired", "lev_i", "lev_j", "Bij", "lower_off", "q", "wq",
                                              "wphi_line", "b", "nstar", "pop_ratio", "inv_df", "vproj", "e_kT",
                                              "adamp", "xhat", "ehat", "split_opac1", "split_opac2",
                                              "split_chi1", "split_chi2", "split_id", "split_chi_id", "wphi")])


class BfRatesArgs(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in ("ncont", "nline", "nf", "ncell", "nk", "precond")]
                + [("cut_x", C.c_double), ("sf", C.c_int64), ("sc", C.c_int64)]
                + [(n, C.c_void_p) for n in ("iblue", "ired", "lev_i", "lev_j", "active", "ev", "wnu", "alpha",
                                              "twohnu3_c2", "e_kT", "nstar", "b", "Jnu", "beta", "PsiJ",
                                              "gamma_ij", "gamma_ji")])


BF_DTYPES = dict(iblue=torch.int32, ired=torch.int32, lev_i=torch.int32, lev_j=torch.int32, active=torch.int32,
                 ev=torch.float64, wnu=torch.float64, alpha=torch.float32, twohnu3_c2=torch.float32,
                 e_kT=torch.float64, nstar=torch.float64, b=torch.float64, Jnu=torch.float32, beta=torch.float32,
                 PsiJ=torch.float32)


def _ali_lib():
    L = lib()
    if not getattr(L, "_ali_ok", False):
        L.m3drt_bb_rates.argtypes = [C.c_void_p, C.POINTER(BbRatesArgs)]
        L.m3drt_bb_opacity.argtypes = [C.c_void_p, C.POINTER(BbOpacityArgs)]
        L.m3drt_bf_rates.argtypes = [C.c_void_p, C.POINTER(BfRatesArgs)]
        for n in ("m3drt_bb_rates", "m3drt_bb_opacity", "m3drt_bf_rates"):
            getattr(L, n).restype = C.c_int
        L._ali_ok = True
    return L


def _check_dev(d, dtypes, what):
    for k, t in d.items():
        if t is not None and (t.dtype != dtypes[k] or not t.is_cuda or not t.is_contiguous()):
            raise M3drtError(f"{what}: {k} must be a contiguous {dtypes[k]} device tensor")


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _strides(layout, nf, ncell):
    if layout == "native":
        return ncell, 1
    if layout == "reference":
        return 1, nf
    raise M3drtError(f"unknown layout {layout!r}")


def _bb_rates(self, d, wmu, wtot, gamma_ij, gamma_ji, precond=False, own_norm=False, largest=False,
              layout="native", split=None, split_offset=0):
    """bound_bound_rates of one direction on the device.  d: dict of device tensors
    (BB_FIELDS, dtypes BB_DTYPES); spectra are (nf, ncell) in the native layout or
    (ncell, nf) in the reference layout; gamma_ij, gamma_ji (nline, ncell) float32 are
    accumulated in place.  With largest and precond, split holds SPLIT_FIELDS (psi
    flags when it holds ntrans, psi_flags_in and psi_flags_out, which is OR-ed into)."""
    L = _ali_lib()
    sp = split or {}
    _check_dev(d, BB_DTYPES, "bb_rates")
    _check_dev(sp, BB_DTYPES, "bb_rates")
    for t in (gamma_ij, gamma_ji):
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise M3drtError("bb_rates: gamma_ij/gamma_ji are contiguous float32 device tensors")
    nline, nf = d["q"].shape
    nk, ncell = d["b"].shape
    sf, sc = _strides(layout, nf, ncell)
    flags = ((M3DRT_RATES_PRECOND if precond else 0) | (M3DRT_RATES_OWN_NORM if own_norm else 0)
             | (M3DRT_RATES_LARGEST if largest else 0) | (M3DRT_RATES_SCAT if sp.get("chi_scat") is not None else 0)
             | (M3DRT_RATES_FLAGS if sp.get("psi_flags_in") is not None else 0))
    npw = int(sp["psi_flags_in"].shape[0]) if sp.get("psi_flags_in") is not None else 0
    a = BbRatesArgs(nline, nf, ncell, nk, flags, wmu, wtot, sf, sc, *[_p(d.get(k)) for k in BB_FIELDS],
                    _p(gamma_ij), _p(gamma_ji), split_offset, npw, *[_p(sp.get(k)) for k in SPLIT_FIELDS])
    self._bind_stream()
    _check(L.m3drt_bb_rates(self._h, C.byref(a)), "m3drt_bb_rates")


def _bb_opacity(self, d, xhat, ehat, wmu=0.0, lte=False, split=None, split_offset=0, wphi_line=None, wphi=None,
                layout="native"):
    """add_bound_bound_opacities of one direction on the device: xhat, ehat (and the
    split tracking arrays, and wphi) are updated in place."""
    L = _ali_lib()
    sp = split or {}
    _check_dev(d, BB_DTYPES, "bb_opacity")
    _check_dev(dict(sp, xhat=xhat, ehat=ehat, wphi=wphi, wphi_line=wphi_line), BB_DTYPES, "bb_opacity")
    nline, nf = d["q"].shape
    nk, ncell = d["nstar"].shape
    sf, sc = _strides(layout, nf, ncell)
    flags = (M3DRT_BBOP_LTE if lte else 0) | (M3DRT_BBOP_SPLIT if split is not None else 0)
    a = BbOpacityArgs(nline, nf, ncell, nk, flags, split_offset, wmu, sf, sc,
                      *[_p(d.get(k)) for k in ("iblue", "ired", "lev_i", "lev_j", "Bij", "lower_off", "q", "wq")],
                      _p(wphi_line),
                      *[_p(d.get(k)) for k in ("b", "nstar", "pop_ratio", "inv_df", "vproj", "e_kT", "adamp")],
                      _p(xhat), _p(ehat),
                      *[_p(sp.get(k)) for k in ("split_opac1", "split_opac2", "split_chi1", "split_chi2", "split_id",
                                                "split_chi_id")],
                      _p(wphi))
    self._bind_stream()
    _check(L.m3drt_bb_opacity(self._h, C.byref(a)), "m3drt_bb_opacity")


class WindowArgs(C.Structure):
    _fields_ = [("flags", C.c_int32), ("maser_chi_floor", C.c_double)] + \
               [(n, C.c_void_p) for n in ("chi_cont", "Bnu", "beta", "n_floored")]


def _window_opacity(self, d, chi_cont, Bnu, beta, xhat, ehat, split=None, split_offset=0, lte=False,
                    suppress_masers=False, maser_chi_floor=1e-4, n_floored=None, layout="native"):
    """One direction's opacity pass over a window in one kernel (m3drt_window_opacity):
    the continuum, convert_to_work_units, the split tracking seeded with the continuum,
    the lines with add_largest, and work_to_transfer_units.  xhat receives chi and
    ehat S; split (the six tracking arrays) is overwritten."""
    L = _ali_lib()
    if not getattr(L, "_win_ok", False):
        L.m3drt_window_opacity.argtypes = [C.c_void_p, C.POINTER(BbOpacityArgs), C.POINTER(WindowArgs)]
        L.m3drt_window_opacity.restype = C.c_int
        L._win_ok = True
    sp = split or {}
    _check_dev(d, BB_DTYPES, "window_opacity")
    _check_dev(dict(sp, xhat=xhat, ehat=ehat), BB_DTYPES, "window_opacity")
    _f32_like(xhat, chi_cont, Bnu, beta)
    nline, nf = d["q"].shape
    nk, ncell = d["nstar"].shape
    sf, sc = _strides(layout, nf, ncell)
    flags = (M3DRT_BBOP_LTE if lte else 0) | (M3DRT_BBOP_SPLIT if split is not None else 0)
    a = BbOpacityArgs(nline, nf, ncell, nk, flags, split_offset, 0.0, sf, sc,
                      *[_p(d.get(k)) for k in ("iblue", "ired", "lev_i", "lev_j", "Bij", "lower_off", "q")], None,
                      None, *[_p(d.get(k)) for k in ("b", "nstar", "pop_ratio", "inv_df", "vproj", "e_kT", "adamp")],
                      _p(xhat), _p(ehat),
                      *[_p(sp.get(k)) for k in ("split_opac1", "split_opac2", "split_chi1", "split_chi2", "split_id",
                                                "split_chi_id")],
                      None)
    w = WindowArgs(M3DRT_UNITS_SUPPRESS_MASERS if suppress_masers else 0, maser_chi_floor, _p(chi_cont), _p(Bnu),
                   _p(beta), _p(n_floored))
    self._bind_stream()
    _check(L.m3drt_window_opacity(self._h, C.byref(a), C.byref(w)), "m3drt_window_opacity")


def _bf_rates(self, d, nline, gamma_ij, gamma_ji, precond=False, cut_x=70.0, layout="native"):
    """bound_free_rates on the device; d: BF_DTYPES tensors, spectra (nf, ncell)
    native or (ncell, nf) reference; gamma arrays (nline + ncont, ncell)."""
    L = _ali_lib()
    _check_dev(d, BF_DTYPES, "bf_rates")
    ncont, nf = d["alpha"].shape
    nk, ncell = d["nstar"].shape
    sf, sc = _strides(layout, nf, ncell)
    a = BfRatesArgs(ncont, nline, nf, ncell, nk, int(precond), cut_x, sf, sc,
                    *[_p(d.get(k)) for k in ("iblue", "ired", "lev_i", "lev_j", "active", "ev", "wnu", "alpha",
                                             "twohnu3_c2", "e_kT", "nstar", "b", "Jnu", "beta", "PsiJ")],
                    _p(gamma_ij), _p(gamma_ji))
    self._bind_stream()
    _check(L.m3drt_bf_rates(self._h, C.byref(a)), "m3drt_bf_rates")


Context.bb_rates = _bb_rates
Context.bb_opacity = _bb_opacity
Context.window_opacity = _window_opacity
Context.bf_rates = _bf_rates


# ------------------------------------------------------------------ working units

M3DRT_UNITS_SPLIT, M3DRT_UNITS_SCAT, M3DRT_UNITS_SUPPRESS_MASERS = 1, 2, 4


class UnitsArgs(C.Structure):
    _fields_ = [("n", C.c_int64), ("flags", C.c_int32), ("maser_chi_floor", C.c_double)] + \
               [(n, C.c_void_p) for n in ("work1", "work2", "beta", "split_opac1", "split_opac2", "chi_scat", "dJnu",
                                          "n_floored")]


def _units_lib():
    L = lib()
    if not getattr(L, "_units_ok", False):
        L.m3drt_work_to_transfer.argtypes = [C.c_void_p, C.POINTER(UnitsArgs)]
        L.m3drt_work_to_transfer.restype = C.c_int
        L.m3drt_convert_to_work_units.argtypes = [C.c_void_p, C.c_int64] + [C.c_void_p] * 4
        L.m3drt_convert_to_work_units.restype = C.c_int
        L._units_ok = True
    return L


def _f32_like(ref, *ts):
    for t in ts:
        if t is not None and (t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous()
                              or t.numel() != ref.numel()):
            raise M3drtError("working-unit arrays are contiguous float32 device tensors of one size")


def _work_to_transfer(self, work1, work2, beta, split_opac1=None, split_opac2=None, chi_scat=None, dJnu=None,
                      suppress_masers=False, maser_chi_floor=1e-4, n_floored=None):
    """work_to_transfer_units in place: work1 xhat -> chi, work2 ehat -> S.
    n_floored: optional int64 device tensor of one element, incremented."""
    L = _units_lib()
    _f32_like(work1, work2, beta, split_opac1, split_opac2, chi_scat, dJnu)
    if n_floored is not None and (n_floored.dtype != torch.int64 or not n_floored.is_cuda):
        raise M3drtError("n_floored is an int64 device tensor")
    flags = ((M3DRT_UNITS_SPLIT if split_opac1 is not None else 0) | (M3DRT_UNITS_SCAT if chi_scat is not None else 0)
             | (M3DRT_UNITS_SUPPRESS_MASERS if suppress_masers else 0))
    a = UnitsArgs(work1.numel(), flags, maser_chi_floor, *[_p(t) for t in (work1, work2, beta, split_opac1,
                                                                           split_opac2, chi_scat, dJnu, n_floored)])
    self._bind_stream()
    _check(L.m3drt_work_to_transfer(self._h, C.byref(a)), "m3drt_work_to_transfer")


def _convert_to_work_units(self, chi_cont, xhat, ehat, beta):
    """convert_to_work_units in place: xhat = chi_cont + ehat, ehat /= beta."""
    L = _units_lib()
    _f32_like(chi_cont, xhat, ehat, beta)
    self._bind_stream()
    _check(L.m3drt_convert_to_work_units(self._h, chi_cont.numel(), _p(chi_cont), _p(xhat), _p(ehat), _p(beta)),
           "m3drt_convert_to_work_units")


Context.work_to_transfer = _work_to_transfer
Context.convert_to_work_units = _convert_to_work_units


# ------------------------------------------------------------------ eta-step flags

M3DRT_ETA_GHOST = 1


class EtaArgs(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nf", "nx", "ny", "nz", "dir", "flags", "rx", "ry")] + \
               [(n, C.c_void_p) for n in ("chi", "S", "beta", "ghost_chi", "ghost_S", "src", "mask", "ghost_src")]


def eta_spread(rays, ds, ds_upstream=None, columns=False):
    """eta_rx, eta_ry of update_eta_mask (SRC/rt_patch_mod.f90:3690-3706): the
    horizontal reach of the upwind point over the slab's directions, plus one."""
    if columns:
        return 0, 0
    dz = max(ds[2], ds_upstream[2]) if ds_upstream is not None else ds[2]
    sx = max(abs(v[0] / v[2]) for v in rays)
    sy = max(abs(v[1] / v[2]) for v in rays)
    return int(np.floor(sx * dz / ds[0])) + 1, int(np.floor(sy * dz / ds[1])) + 1


def _update_eta_mask(self, chi, S, beta, src, mask, *, dir, rx, ry, ghost_chi=None, ghost_S=None,
                     ghost_src=None):
    """update_eta_mask on dense flags in the native layout: chi, S, beta float32
    (nz, nf, ny, nx); src, mask uint8 of that shape; the ghost plane (nf, ny, nx)."""
    L = lib()
    if not getattr(L, "_eta_ok", False):
        L.m3drt_update_eta_mask.argtypes = [C.c_void_p, C.POINTER(EtaArgs)]
        L.m3drt_update_eta_mask.restype = C.c_int
        L._eta_ok = True
    nz, nf, ny, nx = chi.shape
    for t in (chi, S, beta):
        if t.dtype != torch.float32 or tuple(t.shape) != tuple(chi.shape) or not t.is_cuda or not t.is_contiguous():
            raise M3drtError("eta mask: chi, S, beta are float32 (nz, nf, ny, nx) device tensors")
    for t in (src, mask):
        if t.dtype != torch.uint8 or tuple(t.shape) != tuple(chi.shape) or not t.is_cuda or not t.is_contiguous():
            raise M3drtError("eta mask: src, mask are uint8 (nz, nf, ny, nx) device tensors")
    ghost = ghost_chi is not None
    if ghost:
        for t in (ghost_chi, ghost_S, ghost_src):
            if t is None or t.numel() != nf * ny * nx or not t.is_cuda or not t.is_contiguous():
                raise M3drtError("eta mask: the ghost plane is (nf, ny, nx) chi, S (float32) and src (uint8)")
    a = EtaArgs(nf, nx, ny, nz, int(dir), M3DRT_ETA_GHOST if ghost else 0, rx, ry,
                *[_p(t) for t in (chi, S, beta, ghost_chi, ghost_S, src, mask, ghost_src)])
    self._bind_stream()
    _check(L.m3drt_update_eta_mask(self._h, C.byref(a)), "m3drt_update_eta_mask")


Context.update_eta_mask = _update_eta_mask


M3DRT_BG_SCAT_ALI, M3DRT_BG_J_TO_B = 1, 2


class BackgroundArgs(C.Structure):
    _fields_ = [("n", C.c_int64), ("flags", C.c_int32), ("wscat", C.c_float), ("scat_op_max", C.c_double)] + \
               [(n, C.c_void_p) for n in ("J", "B", "chi_cont", "sigma", "PsiJ", "Jali_old", "Jali", "eta", "dJnu",
                                          "n_neg", "diag")]


def _background_emission(self, J, B, chi_cont, sigma, eta, *, PsiJ=None, Jali_old=None, Jali=None, dJnu=None,
                         j_to_b=False, wscat=1.0, scat_op_max=0.9, n_neg=None, diag=None):
    """lte_background_emission in place over a window's flat float32 arrays (J in:
    J_FS, out: J - B); n_neg int64[1] and diag float64[3] device counters, optional."""
    L = lib()
    if not getattr(L, "_bg_ok", False):
        L.m3drt_background_emission.argtypes = [C.c_void_p, C.POINTER(BackgroundArgs)]
        L.m3drt_background_emission.restype = C.c_int
        L._bg_ok = True
    _f32_like(J, B, chi_cont, sigma, eta, PsiJ, Jali_old, Jali, dJnu)
    flags = (M3DRT_BG_SCAT_ALI if PsiJ is not None else 0) | (M3DRT_BG_J_TO_B if j_to_b else 0)
    a = BackgroundArgs(J.numel(), flags, wscat, scat_op_max,
                       *[_p(t) for t in (J, B, chi_cont, sigma, PsiJ, Jali_old, Jali, eta, dJnu, n_neg, diag)])
    self._bind_stream()
    _check(L.m3drt_background_emission(self._h, C.byref(a)), "m3drt_background_emission")


Context.background_emission = _background_emission


# ------------------------------------------------------------------ batched 1D synthesis

B1D_MAXMU = 8


class B1dBatch(C.Structure):
    _fields_ = [("natm", C.c_int32), ("reserved", C.c_int32)] + [(n, C.c_void_p) for n in (
        "nd", "nf", "nmu", "nline", "d_off", "l_off", "ld_off", "mu_off", "f_off", "i_off", "lam0", "dlam",
        "dtau", "hckT", "ivth", "cont_a", "cont_b", "line_off", "line_cl", "line_rec", "mu", "wflux")] + \
        [("lam_ref", C.c_double), ("I_top", C.c_void_p), ("flux", C.c_void_p)]


class B1dTile(C.Structure):
    _fields_ = [("fpl", C.c_int32), ("threads", C.c_int32)]


# device dtypes of the batch's arrays, by field
B1D_DTYPES = dict(nd=torch.int32, nf=torch.int32, nmu=torch.int32, nline=torch.int32, d_off=torch.int64,
                  l_off=torch.int64, ld_off=torch.int64, mu_off=torch.int64, f_off=torch.int64, i_off=torch.int64,
                  lam0=torch.float64, dlam=torch.float64, dtau=torch.float32, hckT=torch.float32,
                  ivth=torch.float32, cont_a=torch.float32, cont_b=torch.float32, line_off=torch.float32,
                  line_cl=torch.float32, line_rec=torch.float32, mu=torch.float32, wflux=torch.float32,
                  I_top=torch.float32, flux=torch.float32)


def _b1d_lib():
    L = lib()
    if not getattr(L, "_b1d_ok", False):
        vp = C.c_void_p
        L.m3drt_b1d_synth_plain.argtypes = [vp, C.POINTER(B1dBatch), C.POINTER(B1dTile)]
        L.m3drt_b1d_plan_create.argtypes = [vp, C.POINTER(B1dBatch), C.POINTER(B1dTile), C.POINTER(vp)]
        L.m3drt_b1d_plan_run.argtypes = [vp]
        L.m3drt_b1d_plan_info.argtypes = [vp, C.POINTER(B1dTile), C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.m3drt_b1d_plan_destroy.argtypes = [vp]
        for n in ("m3drt_b1d_synth_plain", "m3drt_b1d_plan_create", "m3drt_b1d_plan_run", "m3drt_b1d_plan_info",
                  "m3drt_b1d_plan_destroy"):
            getattr(L, n).restype = C.c_int
        L._b1d_ok = True
    return L


def b1d_check_sizes(arrays: dict):
    """Every flat array long enough for the atmospheres' counts and offsets (the
    library takes plain pointers and cannot check their extents)."""
    h = {k: arrays[k].cpu().numpy().astype(np.int64) for k in
         ("nd", "nf", "nmu", "nline", "d_off", "l_off", "ld_off", "mu_off", "f_off", "i_off")}
    n = len(h["nd"])
    for k in ("nf", "nmu", "nline", "d_off", "l_off", "ld_off", "mu_off", "f_off", "i_off", "lam0", "dlam"):
        if arrays[k].numel() != n:
            raise ValueError(f"b1d: {k} must have one entry per atmosphere")
    if n == 0:
        return
    if (h["nd"] < 2).any() or (h["nf"] < 1).any() or (h["nmu"] < 1).any() or (h["nmu"] > B1D_MAXMU).any() or \
            (h["nline"] < 0).any():
        raise ValueError("b1d: every atmosphere needs nd >= 2, nf >= 1, 1 <= nmu <= 8, nline >= 0")
    for k in ("d_off", "l_off", "ld_off", "mu_off", "f_off", "i_off"):
        if (h[k] < 0).any():
            raise ValueError(f"b1d: negative {k}")
    need = dict(dtau=h["d_off"] + h["nd"], hckT=h["d_off"] + h["nd"], ivth=h["d_off"] + h["nd"],
                cont_a=h["d_off"] + h["nd"], cont_b=h["d_off"] + h["nd"], line_off=h["l_off"] + h["nline"],
                line_cl=h["l_off"] + h["nline"], line_rec=4 * (h["ld_off"] + h["nd"] * h["nline"]),
                mu=h["mu_off"] + h["nmu"], wflux=h["mu_off"] + h["nmu"], flux=h["f_off"] + h["nf"],
                I_top=h["i_off"] + h["nmu"] * h["nf"])
    for k, v in need.items():
        if arrays[k].numel() < int(v.max()):
            raise ValueError(f"b1d: {k} has {arrays[k].numel()} entries, the offsets need {int(v.max())}")


def _b1d_struct(arrays: dict, lam_ref: float) -> B1dBatch:
    """The m3drt_b1d_batch of a dict of contiguous device tensors (B1D_DTYPES)."""
    for k, dt in B1D_DTYPES.items():
        t = arrays[k]
        if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"b1d: {k} must be a contiguous {dt} device tensor")
    b1d_check_sizes(arrays)
    b = B1dBatch()
    b.natm = arrays["nd"].numel()
    for k in B1D_DTYPES:
        setattr(b, k, arrays[k].data_ptr())
    b.lam_ref = lam_ref
    return b


def _b1d_tile(tile):
    return None if tile is None else C.byref(B1dTile(*tile))


def _b1d_synth_plain(self, arrays, lam_ref, tile=None):
    """One launch per atmosphere (m3drt_b1d_synth_plain); tile = (fpl, threads)."""
    L = _b1d_lib()
    b = _b1d_struct(arrays, lam_ref)
    self._bind_stream()
    _check(L.m3drt_b1d_synth_plain(self._h, C.byref(b), _b1d_tile(tile)), "m3drt_b1d_synth_plain")


class B1dPlan:
    """m3drt_b1d_plan: the persistent, graph-captured synthesis of one batch; tile None
    lets the autotuner choose.  Keeps the batch's tensors alive."""

    def __init__(self, ctx, arrays, lam_ref, tile=None):
        self.ctx, self.arrays = ctx, arrays
        self._b = _b1d_struct(arrays, lam_ref)
        L = _b1d_lib()
        h = C.c_void_p()
        ctx._bind_stream()
        _check(L.m3drt_b1d_plan_create(ctx._h, C.byref(self._b), _b1d_tile(tile), C.byref(h)),
               "m3drt_b1d_plan_create")
        self._h = h

    def run(self):
        self.ctx._bind_stream()
        _check(_b1d_lib().m3drt_b1d_plan_run(self._h), "m3drt_b1d_plan_run")

    def info(self):
        """(tile (fpl, threads), the autotuner's times in ms per candidate)."""
        t = B1dTile()
        times = (C.c_double * 6)()
        n = C.c_int32()
        _check(_b1d_lib().m3drt_b1d_plan_info(self._h, C.byref(t), times, C.byref(n)), "m3drt_b1d_plan_info")
        return (t.fpl, t.threads), list(times)[:n.value]

    def close(self):
        if self._h:
            _b1d_lib().m3drt_b1d_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


Context.b1d_synth_plain = _b1d_synth_plain
Context.b1d_plan = lambda self, arrays, lam_ref, tile=None: B1dPlan(self, arrays, lam_ref, tile)


# ------------------------------------------------------------------ rotation for the output pass

class RotateArgs(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nf", "nx", "ny", "nz", "nk", "reserved")] + \
               [(n, C.c_int64) for n in ("cs_f", "cs_x", "cs_y", "cs_z", "cs_k")] + \
               [("vec", C.c_double * 3), ("ds", C.c_double * 3), ("bot2pivot", C.c_double),
                ("ray_offset", C.c_double * 2), ("in", C.c_void_p), ("out", C.c_void_p)]


def _rotate_array(self, x, out, vec, ds, bot2pivot, ray_offset=(0.0, 0.0), layout="reference"):
    """rotate_array of the cubes x into out (distinct float32 device tensors of one
    shape): layout "reference" is (k, z, y, x, f) in C order -- work(f, x, y, z, k);
    "native" is (k, z, f, y, x)."""
    L = lib()
    if not getattr(L, "_rot_ok", False):
        L.m3drt_rotate_array.argtypes = [C.c_void_p, C.POINTER(RotateArgs)]
        L.m3drt_rotate_array.restype = C.c_int
        L._rot_ok = True
    for t in (x, out):
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.dim() != 5:
            raise ValueError("rotate_array: contiguous 5-D float32 device tensors")
    if x.shape != out.shape or x.data_ptr() == out.data_ptr():
        raise ValueError("rotate_array: out must be a distinct tensor of x's shape")
    if layout == "reference":
        nk, nz, ny, nx, nf = x.shape
        cs = (1, nf, nf * nx, nf * nx * ny, nf * nx * ny * nz)
    else:
        nk, nz, nf, ny, nx = x.shape
        cs = (nx * ny, 1, nx, nx * ny * nf, nx * ny * nf * nz)
    a = RotateArgs(nf, nx, ny, nz, nk, 0, *cs, (C.c_double * 3)(*vec), (C.c_double * 3)(*ds), bot2pivot,
                   (C.c_double * 2)(*ray_offset), x.data_ptr(), out.data_ptr())
    self._bind_stream()
    _check(L.m3drt_rotate_array(self._h, C.byref(a)), "m3drt_rotate_array")


Context.rotate_array = _rotate_array



def _weights(self, t, which):
    L = lib()
    if not getattr(L, "_w_ok", False):
        for n in ("m3drt_linear_weights", "m3drt_parab_moments"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_int64] + [C.c_void_p] * 5
            getattr(L, n).restype = C.c_int
        L._w_ok = True
    if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
        raise M3drtError("weights: t is a contiguous float32 device tensor")
    out = [torch.empty_like(t) for _ in range(4)]
    self._bind_stream()
    name = "m3drt_linear_weights" if which == "linear" else "m3drt_parab_moments"
    _check(getattr(L, name)(self._h, t.numel(), _p(t), *[_p(o) for o in out]), name)
    return tuple(out)


def _linear_weights(self, t):
    """linear_weights of the kernels elementwise: (E, W, Psi, Esc)."""
    return _weights(self, t, "linear")


def _parab_moments(self, t):
    """parab_moments of the kernels elementwise: (E, U0, U1, q)."""
    return _weights(self, t, "parab")


Context.linear_weights = _linear_weights
Context.parab_moments = _parab_moments


class QuadratureArgs(C.Structure):
    _fields_ = [("scheme", C.c_char_p), ("ntheta", C.c_int32), ("nphi", C.c_int32), ("jnu_weights", C.c_int32),
                ("one_dimensional", C.c_int32), ("mu", C.c_void_p), ("ds", C.c_double * 3)]


def quadrature(scheme, ntheta=0, nphi=1, jnu_weights=True, one_dimensional=False, mu=None, ds=(1.0, 1.0, 1.0)):
    """setup_quadrature's ray set: (vec (n, 3), wts (n)) as float64 numpy arrays."""
    L = lib()
    if not getattr(L, "_quad_ok", False):
        L.m3drt_quadrature.argtypes = [C.POINTER(QuadratureArgs), C.c_int32, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_int32)]
        L.m3drt_quadrature.restype = C.c_int
        L._quad_ok = True
    m = None if mu is None else np.ascontiguousarray(mu, np.float64)
    if m is not None:
        ntheta = m.size
    q = QuadratureArgs(scheme.encode(), int(ntheta), int(nphi), int(bool(jnu_weights)), int(bool(one_dimensional)),
                       None if m is None else m.ctypes.data_as(C.c_void_p), (C.c_double * 3)(*ds))
    n = C.c_int32(0)
    st = L.m3drt_quadrature(C.byref(q), 0, None, None, C.byref(n))
    if st not in (M3DRT_OK, 4):
        _check(st, "m3drt_quadrature")
    vec = np.zeros((n.value, 3))
    w = np.zeros(n.value)
    _check(L.m3drt_quadrature(C.byref(q), n.value, vec.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
                              C.byref(n)), "m3drt_quadrature")
    return vec, w


class BackgroundLteArgs(C.Structure):
    _fields_ = [("nf", C.c_int32), ("ncell", C.c_int64), ("nu", C.c_void_p), ("temp", C.c_void_p),
                ("beta", C.c_void_p), ("Bnu", C.c_void_p), ("sf", C.c_int64), ("sc", C.c_int64)]


class BoundaryArgs(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in ("kind", "nf", "nx", "ny")]
                + [(n, C.c_int64) for n in ("ps_f", "ps_x", "ps_y")]
                + [(n, C.c_void_p) for n in ("chi1", "S1", "chi2", "S2", "g500", "chi500_1", "T_ghost", "nu")]
                + [("ds3", C.c_double)]
                + [(n, C.c_void_p) for n in ("exit_I", "mem_chi", "mem_S", "mem_I", "ghost_chi", "ghost_S",
                                              "ghost_I")])


M3DRT_GHOST = {"top": 0, "lte": 1, "pchip": 2, "upstream": 3}


def _bnd_lib():
    L = lib()
    if not getattr(L, "_bnd_ok", False):
        L.m3drt_lte_background.argtypes = [C.c_void_p, C.POINTER(BackgroundLteArgs)]
        L.m3drt_lte_background.restype = C.c_int
        for n in ("m3drt_set_upstream_boundary", "m3drt_set_downstream_boundary"):
            getattr(L, n).argtypes = [C.c_void_p, C.POINTER(BoundaryArgs)]
            getattr(L, n).restype = C.c_int
        L._bnd_ok = True
    return L


def _lte_background(self, nu, temp, beta, Bnu):
    """beta and Bnu (nf, ncell) float32 device tensors from nu (nf) and temp (ncell)
    float64 device tensors (lte_background_opacities' factors)."""
    L = _bnd_lib()
    for t, dt in ((nu, torch.float64), (temp, torch.float64), (beta, torch.float32), (Bnu, torch.float32)):
        if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
            raise M3drtError("lte_background: nu, temp float64, beta, Bnu float32 contiguous device tensors")
    nf, nc = nu.numel(), temp.numel()
    if beta.numel() != nf * nc or Bnu.numel() != nf * nc:
        raise M3drtError("lte_background: beta and Bnu hold nf * ncell values")
    a = BackgroundLteArgs(nf, nc, _p(nu), _p(temp), _p(beta), _p(Bnu), nc, 1)
    self._bind_stream()
    _check(L.m3drt_lte_background(self._h, C.byref(a)), "m3drt_lte_background")


def _boundary_args(kind, ghost, nf, ny, nx, **kw):
    a = BoundaryArgs()
    a.kind = M3DRT_GHOST[kind] if isinstance(kind, str) else int(kind)
    a.nf, a.nx, a.ny = nf, nx, ny
    a.ps_f, a.ps_x, a.ps_y = nx * ny, 1, nx
    a.ds3 = float(kw.pop("ds3", 0.0))
    for k, t in kw.items():
        if t is not None:
            if not t.is_cuda or not t.is_contiguous():
                raise M3drtError(f"boundary: {k} is a contiguous device tensor")
            setattr(a, k, t.data_ptr())
    if ghost is not None:
        a.ghost_chi, a.ghost_S, a.ghost_I = [t.data_ptr() for t in ghost]
    return a


def _set_upstream_boundary(self, kind, ghost, *, chi1=None, S1=None, chi2=None, S2=None, g500=None, chi500_1=None,
                           T_ghost=None, nu=None, ds3=0.0, mem=None):
    """set_upstream_boundary into ghost = (chi, S, I), native (nf, ny, nx) float32
    planes: kind "top", "lte", "pchip" or "upstream" (mem = (chi, S, I))."""
    L = _bnd_lib()
    nf, ny, nx = ghost[0].shape
    mc, ms, mi = mem if mem is not None else (None, None, None)
    a = _boundary_args(kind, ghost, nf, ny, nx, chi1=chi1, S1=S1, chi2=chi2, S2=S2, g500=g500, chi500_1=chi500_1,
                       T_ghost=T_ghost, nu=nu, ds3=ds3, mem_chi=mc, mem_S=ms, mem_I=mi)
    self._bind_stream()
    _check(L.m3drt_set_upstream_boundary(self._h, C.byref(a)), "m3drt_set_upstream_boundary")


def _set_downstream_boundary(self, chi_exit, S_exit, I_exit, mem):
    """set_downstream_boundary: the exit plane (chi, S, I) into mem = (chi, S, I)."""
    L = _bnd_lib()
    nf, ny, nx = chi_exit.shape
    a = _boundary_args("top", None, nf, ny, nx, chi1=chi_exit, S1=S_exit, exit_I=I_exit, mem_chi=mem[0],
                       mem_S=mem[1], mem_I=mem[2])
    self._bind_stream()
    _check(L.m3drt_set_downstream_boundary(self._h, C.byref(a)), "m3drt_set_downstream_boundary")


Context.lte_background = _lte_background
Context.set_upstream_boundary = _set_upstream_boundary
Context.set_downstream_boundary = _set_downstream_boundary


class OptabTable(C.Structure):
    _fields_ = [("nnu", C.c_int32), ("ntemp", C.c_int32), ("nne", C.c_int32), ("mu", C.c_double * 2),
                ("rot", C.c_double * 4), ("c0", C.c_double * 2), ("dc", C.c_double * 2), ("lam", C.c_void_p),
                ("val", C.c_void_p), ("scat", C.c_void_p)]


class OptabQuery(C.Structure):
    _fields_ = [("ncell", C.c_int64), ("nf", C.c_int32), ("temp", C.c_void_p), ("ne", C.c_void_p),
                ("chi500", C.c_void_p), ("lam_q", C.c_void_p), ("back", C.c_void_p), ("scat", C.c_void_p),
                ("sf", C.c_int64), ("sc", C.c_int64)]


class OptabPlan:
    """The sampled background table on the device (m3drt_optab_plan_*).  lam (nnu)
    float64, val and scat (nne, ntemp, nnu) float32 device tensors -- the
    reference's (nnu, ntemp, nne) with nnu fastest; mu, rot (2x2, PC directions as
    columns), c0, dc of the PC axes."""

    def __init__(self, ctx, lam, val, scat, mu, rot, c0, dc):
        L = lib()
        if not getattr(L, "_optab_ok", False):
            L.m3drt_optab_plan_create.argtypes = [C.c_void_p, C.POINTER(OptabTable), C.POINTER(C.c_void_p)]
            L.m3drt_optab_slopes.argtypes = [C.c_void_p] * 7
            L.m3drt_optab_interpolate.argtypes = [C.c_void_p, C.POINTER(OptabQuery)]
            L.m3drt_optab_plan_destroy.argtypes = [C.c_void_p]
            for n in ("m3drt_optab_plan_create", "m3drt_optab_slopes", "m3drt_optab_interpolate",
                      "m3drt_optab_plan_destroy"):
                getattr(L, n).restype = C.c_int
            L._optab_ok = True
        for t, dt in ((lam, torch.float64), (val, torch.float32), (scat, torch.float32)):
            if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
                raise M3drtError("optab: lam float64, val and scat float32 contiguous device tensors")
        ne_, nt, nnu = val.shape
        rot = np.asarray(rot, np.float64)
        t = OptabTable(nnu, nt, ne_, (C.c_double * 2)(*mu), (C.c_double * 4)(*rot.ravel(order="F")),
                       (C.c_double * 2)(*c0), (C.c_double * 2)(*dc), lam.data_ptr(), val.data_ptr(), scat.data_ptr())
        self.ctx, self.shape = ctx, val.shape
        self._h = C.c_void_p()
        ctx._bind_stream()
        _check(L.m3drt_optab_plan_create(ctx._h, C.byref(t), C.byref(self._h)), "m3drt_optab_plan_create")

    def slopes(self):
        out = [torch.empty(self.shape, dtype=torch.float32, device="cuda") for _ in range(6)]
        self.ctx._bind_stream()
        _check(lib().m3drt_optab_slopes(self._h, *[_p(o) for o in out]), "m3drt_optab_slopes")
        return out

    def interpolate(self, temp, ne, chi500, lam_q, back, scat):
        """back, scat (nf, ncell) float32 device tensors."""
        nc, nf = temp.numel(), lam_q.numel()
        if back.numel() != nc * nf or scat.numel() != nc * nf:
            raise M3drtError("optab: back and scat hold nf * ncell values")
        q = OptabQuery(nc, nf, _p(temp), _p(ne), _p(chi500), _p(lam_q), _p(back), _p(scat), nc, 1)
        self.ctx._bind_stream()
        _check(lib().m3drt_optab_interpolate(self._h, C.byref(q)), "m3drt_optab_interpolate")

    def close(self):
        if self._h:
            lib().m3drt_optab_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Isotope(C.Structure):
    _fields_ = [("lines", Linelist), ("awgt", C.c_double), ("frac", C.c_double)]


class Species(C.Structure):
    _fields_ = [("niso", C.c_int32), ("iso", C.c_void_p), ("n", C.c_void_p), ("scale", C.c_double),
                ("part", C.c_void_p), ("dep", C.c_void_p), ("nlev", C.c_int32), ("scat_group", C.c_int32)]


class LinelistOpacityArgs(C.Structure):
    _fields_ = ([("nf", C.c_int32), ("nsp", C.c_int32), ("equidistant", C.c_int32), ("ncell", C.c_int64),
                 ("llam", C.c_void_p), ("sp", C.c_void_p)]
                + [(n, C.c_void_p) for n in ("temp", "vmic", "vproj", "e_kT", "ne", "tothi", "stark_pe", "all_chi",
                                              "all_eta", "one_m_eps", "chi_scat", "beta", "chi_lines", "eta_lines")]
                + [("tolerance", C.c_float), ("line_reject_frac", C.c_float)])


def _add_linelist_opacities(self, species, llam, all_chi, all_eta, cells, equidistant=True, one_m_eps=None,
                            chi_scat=None, beta=None, chi_lines=None, eta_lines=None, tolerance=1e-4,
                            line_reject_frac=0.0):
    """add_linelist_opacities for metals and molecules on the device.  species: list
    of dicts -- isotopes: list of (lines dict of device tensors as _line_opacity_equi
    takes, awgt, frac); n, part (ncell) float64 device tensors; scale; dep (nlev,
    ncell) float32 or None; scat_group.  Spectra (all_chi, all_eta, chi_scat, beta,
    chi_lines, eta_lines) are (nf, ncell) float32, one_m_eps (2, nf, ncell); cells:
    temp, vmic, vproj, ekt, ne, tothi, spe (ncell) float32."""
    L = lib()
    if not getattr(L, "_llo_ok", False):
        L.m3drt_add_linelist_opacities.argtypes = [C.c_void_p, C.POINTER(LinelistOpacityArgs)]
        L.m3drt_add_linelist_opacities.restype = C.c_int
        L._llo_ok = True
    nf, ncell = all_chi.shape
    for t in (all_chi, all_eta, chi_scat, beta, chi_lines, eta_lines):
        if t is not None and (t.dtype != torch.float32 or tuple(t.shape) != (nf, ncell) or not t.is_contiguous()):
            raise M3drtError("spectra are contiguous (nf, ncell) float32 device tensors")
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    keep = []
    sps = (Species * max(1, len(species)))()
    for s, sp in enumerate(species):
        isos = (Isotope * max(1, len(sp["isotopes"])))()
        for k, (lines, awgt, frac) in enumerate(sp["isotopes"]):
            nl = int(lines["llam0"].numel())
            sel = _sel_work(lines, nl)
            keep.append(sel)
            isos[k] = Isotope(Linelist(nl, *[p(lines.get(n)) for n in (
                "llam0", "lower_ev", "gpref", "grad", "gw1", "gw2", "gstk", "unsold", "lev_i", "lev_j", "ipart")],
                None, None, None, p(lines.get("reach_in")), p(lines.get("active")), p(sel)), float(awgt),
                float(frac))
        keep.append(isos)
        dep = sp.get("dep")
        sps[s] = Species(len(sp["isotopes"]), C.cast(isos, C.c_void_p), p(sp["n"]), float(sp.get("scale", 1.0)),
                         p(sp["part"]), p(dep), 0 if dep is None else dep.shape[0], int(sp.get("scat_group", 0)))
    a = LinelistOpacityArgs(nf, len(species), int(bool(equidistant)), ncell, p(llam), C.cast(sps, C.c_void_p),
                            *[p(cells[k]) for k in ("temp", "vmic", "vproj", "ekt", "ne", "tothi", "spe")],
                            p(all_chi), p(all_eta), p(one_m_eps), p(chi_scat), p(beta), p(chi_lines), p(eta_lines),
                            tolerance, line_reject_frac)
    self._bind_stream()
    _check(L.m3drt_add_linelist_opacities(self._h, C.byref(a)), "m3drt_add_linelist_opacities")


Context.add_linelist_opacities = _add_linelist_opacities
