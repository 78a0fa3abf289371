#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden from the reference's own compiled code.

Run in the build container only (it needs the reference tree and amdflang):

    bash oracle/build_ref.sh && python tests/golden/make_golden.py

oracle/_ref/hop_golden.x is the harness of oracle/ref_harness/hop_golden.f90
linked against HOP objects compiled from the reference sources by
oracle/build_ref.sh (amdflang, -O2, -DHOP_STANDALONE, no fast-math).  Each fixture
holds the seeded inputs and the reference's outputs, plus provenance.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "hop_golden.x")
EXE_RT = os.path.join(ROOT, "oracle", "_ref", "rt_golden.x")


def provenance(rt=False):
    fc = subprocess.run(["amdflang", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    if rt:
        return dict(compiler=fc, flags="-cpp -DHOP_STANDALONE -O2 -fPIC -fopenmp",
                    sources="experiments/OpenMULTI3D/SRC/ray_mod.f90, solvers/rt/radau_mod.f90, "
                            "experiments/OpenMULTI3D/HOP/{interpol_mod,optab_mod,...}.f90",
                    harness="oracle/ref_harness/rt_golden.f90")
    return dict(compiler=fc, flags="-cpp -DHOP_STANDALONE -O2 -fPIC -fopenmp",
                sources="experiments/OpenMULTI3D/HOP/{qlg_mod,op_math_mod,linelist_mod,...}.f90, exp_neg.inc",
                harness="oracle/ref_harness/hop_golden.f90")


def run(kind, payload: bytes, exe=None) -> bytes:
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        open(fi, "wb").write(payload)
        subprocess.run([exe or EXE, kind, fi, fo], check=True)
        return open(fo, "rb").read()


ONLY: set = set()


def save(name, rt=False, **arrays):
    if ONLY and name not in ONLY:
        return
    prov = provenance(rt)
    h = hashlib.sha256()
    for k in sorted(arrays):
        if k.startswith("in_"):
            h.update(np.ascontiguousarray(arrays[k]).tobytes())
    path = os.path.join(HERE, name)
    np.savez_compressed(path, input_sha256=h.hexdigest(), **{f"prov_{k}": v for k, v in prov.items()},
                        **arrays)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


def i32(*v):
    return np.array(v, np.int32).tobytes()


def gen_expneg(rng):
    t = np.concatenate([np.linspace(-2, 90, 6001), rng.uniform(-2, 87, 4000),
                        10 ** rng.uniform(-8, 0, 1000) * rng.choice([-1, 1], 1000)]).astype(np.float32)
    out = np.frombuffer(run("expneg", i32(t.size) + t.tobytes()), np.float32)
    save("hop_expneg.npz", in_t=t, E=out[:t.size], u0=out[t.size:])


def gen_qlg(rng):
    n = 20000
    a = (10 ** rng.uniform(-4, 2, n)).astype(np.float32)
    a[:200] = 0.0
    v = np.concatenate([rng.uniform(-12, 12, n // 2), rng.uniform(-300, 300, n // 4),
                        rng.uniform(-1, 1, n - n // 2 - n // 4)]).astype(np.float32)
    K = np.frombuffer(run("qlgbot", i32(n) + a.tobytes() + v.tobytes()), np.float32)
    m = 5000
    ad = 10 ** rng.uniform(-4, 2, m)
    vd = np.concatenate([rng.uniform(-20, 20, m // 2), rng.uniform(-100, 100, m - m // 2)])
    Kr = np.frombuffer(run("qlgref", i32(m) + ad.tobytes() + vd.tobytes()), np.float64)
    # line sweeps on a uniform ln(lambda) grid
    nf, nc = 160, 60
    llam = np.log(5000.0) + np.arange(nf) * 7e-6
    idl = 1.0 / 7e-6
    cases = []
    for c in range(nc):
        D = np.float32(rng.uniform(0.3, 4.0))
        gL = np.float32(D * 10 ** rng.uniform(-4, 1.5))
        S = np.float32(10 ** rng.uniform(-6, -2))
        ref = np.float32(10 ** rng.uniform(-10, -6))
        lnq = np.float32(max(0.0, np.log(S / D / ref)))
        cap = np.float32(3.4e38 if c % 5 else rng.uniform(1, 30))
        llc = llam[0] + rng.uniform(-20, nf + 20) * 7e-6
        pos = (llc - llam[0]) * idl
        cases.append((D, gL, S, ref, lnq, cap, llc, pos, idl))
    payload = i32(nc, nf) + llam.tobytes()
    for (D, gL, S, ref, lnq, cap, llc, pos, idl_) in cases:
        payload += np.array([D, gL, S, ref, lnq, cap], np.float32).tobytes()
        payload += np.array([llc, pos, idl_], np.float64).tobytes()
    out = run("qlgline", payload)
    rec = 4 + 4 + 4 + 4 * nf
    fa, fb, hw, Ks = [], [], [], []
    for c in range(nc):
        r = out[c * rec:(c + 1) * rec]
        fa.append(np.frombuffer(r[0:4], np.int32)[0])
        fb.append(np.frombuffer(r[4:8], np.int32)[0])
        hw.append(np.frombuffer(r[8:12], np.float32)[0])
        Ks.append(np.frombuffer(r[12:], np.float32))
    cs = np.array([c[:6] for c in cases], np.float32)
    cd = np.array([c[6:] for c in cases], np.float64)
    save("hop_qlg.npz", in_a=a, in_v=v, K=K, in_ad=ad, in_vd=vd, Kref=Kr, in_llam=llam,
         in_case_f32=cs, in_case_f64=cd, line_fa=np.array(fa, np.int32), line_fb=np.array(fb, np.int32),
         line_hw=np.array(hw, np.float32), line_K=np.array(Ks, np.float32))


def gen_lineq(rng):
    n, ns = 21, 40
    As, bs, xs = [], [], []
    payload = i32(n, ns)
    for s in range(ns):
        # diagonally dominant rate-matrix-like systems (columns summing to ~0 plus a
        # conservation row), the shape update_populations solves
        A = -rng.uniform(0, 1, (n, n)) * (rng.random((n, n)) < 0.5)
        np.fill_diagonal(A, 0.0)
        np.fill_diagonal(A, -A.sum(axis=0) + rng.uniform(0.1, 1, n))
        A[rng.integers(n)] = rng.uniform(0.5, 2, n)
        b = rng.standard_normal(n)
        As.append(A)
        bs.append(b)
        payload += np.asfortranarray(A).tobytes(order="F") + b.tobytes()
    x = np.frombuffer(run("lineq", payload), np.float64).reshape(ns, n)
    save("hop_lineq.npz", in_A=np.array(As), in_b=np.array(bs), x=x)


def gen_opac(rng, name, nl=40, nf=300, m=(4, 3, 2), flags=0, np_=1, nlev=4, general=False):
    """general: add_opacity on a grid whose steps vary from 0.3 to 1.7 of the mean,
    else add_opacity_equi on an equidistant grid.  flags & 8: a list with reaches
    (has_reach) and active flags: the reaches are those a first run without them
    reports, each scaled by 0.3-1.5 (so some windows are capped and some lines are
    not selected), a small one for lines that run did not reach; one line in six is
    inactive."""
    ncell = int(np.prod(m))
    dl = np.log(5010 / 5000) / (nf - 1)
    if general:
        steps = dl * rng.uniform(0.3, 1.7, nf - 1)
        llam = np.log(5000.0) + np.concatenate([[0.0], np.cumsum(steps)])
    else:
        llam = np.log(5000.0) + np.arange(nf) * dl
    llam0 = np.sort(rng.uniform(llam[0] - 40 * dl, llam[-1] + 40 * dl, nl))
    lower_ev = rng.uniform(0, 5, nl).astype(np.float32)
    gpref = (rng.uniform(-3, 0, nl) * np.log(10) - 37.7).astype(np.float32)
    gpref[:3] += 12.0     # a few very strong lines, whose cores raise the running floor
    grad = (10 ** rng.uniform(-9, -7.5, nl)).astype(np.float32)
    gw1 = (10 ** rng.uniform(-30, -28.5, nl)).astype(np.float32)
    gw2 = rng.uniform(0.3, 0.4, nl).astype(np.float32)
    gstk = (10 ** rng.uniform(-24, -22, nl) * rng.choice([1, -1], nl)).astype(np.float32)
    unsold = (rng.random(nl) < 0.3).astype(np.int32)
    lev_i = rng.integers(0, nlev + 1, nl).astype(np.int32)
    lev_j = rng.integers(0, nlev + 1, nl).astype(np.int32)
    ipart = rng.integers(1, np_ + 1, nl).astype(np.int32)
    F = lambda a: np.asfortranarray(a.astype(np.float32))
    cont = (1e-6 * 10 ** rng.uniform(-0.5, 0.5, (1, ncell)) * (1 + 0.1 * np.linspace(0, 1, nf))[:, None])
    cont = cont.astype(np.float32)
    with np.errstate(divide="ignore"):
        logref = (np.float32(np.log(np.float32(1e-4) / np.float32(0.5641895835))) +
                  np.log(cont)).astype(np.float32)
    chi = (cont * rng.uniform(0, 0.5, (nf, ncell))).astype(np.float32)
    eta = (chi * rng.uniform(0.5, 1.0, (nf, ncell))).astype(np.float32)
    cell = lambda lo, hi: rng.uniform(lo, hi, ncell).astype(np.float32)
    vproj = (rng.uniform(-1, 1, ncell) * 5 * dl).astype(np.float32)
    idf = (10 ** rng.uniform(5.0, 5.5, ncell)).astype(np.float32)
    vb2 = (10 ** rng.uniform(11.5, 12.5, ncell)).astype(np.float32)
    ekt = cell(1.5, 2.5)
    ne = (10 ** rng.uniform(12, 15, ncell)).astype(np.float32)
    tothi = (10 ** rng.uniform(15, 17, ncell)).astype(np.float32)
    spe = (ne * 3.0).astype(np.float32)
    lnbase = rng.uniform(20, 36, (np_, ncell)).astype(np.float32)
    dep = rng.uniform(0.5, 2.0, (nlev, ncell)).astype(np.float32)
    ome = csc = None
    if flags & 4:
        ome = rng.uniform(0, 1, (nf, ncell)).astype(np.float32)
        csc = (cont * 0.1).astype(np.float32)

    def payload_of(fl, reach_in=None, active=None):
        pl = i32(nl, nf, *m, np_, nlev, fl) + llam0.tobytes()
        for a in (lower_ev, gpref, grad, gw1, gw2, gstk):
            pl += a.tobytes()
        for a in (unsold, lev_i, lev_j, ipart):
            pl += a.tobytes()
        if fl & 8:
            pl += reach_in.tobytes() + active.tobytes()
        pl += llam.tobytes()
        for a in (chi, eta, cont, logref):
            pl += a.T.tobytes()       # (nf, ncell) in Fortran order = (ncell, nf) C
        for a in (vproj, idf, vb2, ekt, ne, tothi, spe):
            pl += a.tobytes()
        pl += lnbase.tobytes() + dep.tobytes()   # (ncell, np) Fortran = (np, ncell) C
        if fl & 4:
            pl += ome.T.tobytes() + csc.T.tobytes()
        return pl

    kind = "opac_gen" if general else "opac_equi"
    reach_in = active = None
    n2 = nf * ncell
    if flags & 8:
        first = run(kind, payload_of(flags & ~8))
        r0 = np.frombuffer(first[8 * n2 + 8 * nl:8 * n2 + 12 * nl], np.float32)
        reach_in = np.where(r0 > 0, r0 * rng.uniform(0.3, 1.5, nl), rng.uniform(0, 5 * dl, nl)).astype(np.float32)
        active = (rng.random(nl) >= 1 / 6).astype(np.int32)
    out = run(kind, payload_of(flags, reach_in, active))
    o = np.frombuffer(out, np.uint8)
    chi_o = np.frombuffer(out[:4 * n2], np.float32).reshape(ncell, nf)
    eta_o = np.frombuffer(out[4 * n2:8 * n2], np.float32).reshape(ncell, nf)
    p = 8 * n2
    used = np.frombuffer(out[p:p + 4 * nl], np.int32); p += 4 * nl
    strength = np.frombuffer(out[p:p + 4 * nl], np.float32); p += 4 * nl
    reach = np.frombuffer(out[p:p + 4 * nl], np.float32); p += 4 * nl
    extra = {}
    if flags & 8:
        extra["in_reach_in"] = reach_in
        extra["in_active"] = active
    if flags & 4:
        extra["chi_scat"] = np.frombuffer(out[p:p + 4 * n2], np.float32).reshape(ncell, nf)
        extra["in_one_m_eps"] = ome.T.copy()
        extra["in_chi_scat"] = csc.T.copy()
    save(name, in_flags=np.int32(flags), in_m=np.array(m, np.int32), in_llam0=llam0, in_lower_ev=lower_ev,
         in_gpref=gpref, in_grad=grad, in_gw1=gw1, in_gw2=gw2, in_gstk=gstk, in_unsold=unsold,
         in_lev_i=lev_i, in_lev_j=lev_j, in_ipart=ipart, in_llam=llam, in_chi=chi.T.copy(),
         in_eta=eta.T.copy(), in_cont=cont.T.copy(), in_logref=logref.T.copy(), in_vproj=vproj,
         in_idf=idf, in_vb2=vb2, in_ekt=ekt, in_ne=ne, in_tothi=tothi, in_spe=spe, in_lnbase=lnbase,
         in_dep=dep, chi=chi_o, eta=eta_o, used=used, strength=strength, reach=reach, **extra)


QUAD_DS = ((50e5, 50e5, 10e5), (20e5, 30e5, 10e5))


def quad_requests():
    """(scheme, ntheta, nphi, Jnu_weights, one_dimensional, ds, mu) of every set the
    passes can ask setup_quadrature for."""
    req = []
    for jnu in (1, 0):
        for n in range(1, 11):
            for nphi in (1, 4):
                req.append(("radau", n, nphi, jnu, 0, QUAD_DS[0], ()))
        req.append(("radau", 5, 4, jnu, 1, QUAD_DS[0], ()))
        for n in range(1, 7):
            req.append(("lobatto", n, 4, jnu, 0, QUAD_DS[0], ()))
            req.append(("gauss", n, 6, jnu, 0, QUAD_DS[0], ()))
        req.append(("gauss", 3, 6, jnu, 1, QUAD_DS[0], ()))
        for c in ("disk_center", "set_a2", "set_a4", "set_a6", "set_a8", "set_b4", "set_b6", "set_b8"):
            req.append((c, 0, 0, jnu, 0, QUAD_DS[0], ()))
    for mu in ((0.7,), (1.0, 0.5, 0.2), (0.9, 0.6, 0.3, 0.1), (0.2, 1.0, 0.55, 0.8, 0.35)):
        req.append(("custom", len(mu), 4, 0, 0, QUAD_DS[0], mu))
    for g in ("grid+", "gridx", "grid*"):
        for ds in QUAD_DS:
            req.append((g, 0, 4, 0, 0, ds, ()))
    return req


def gen_quad():
    req = quad_requests()
    payload = i32(len(req))
    for (sch, nt, nphi, jnu, one_d, ds, mu) in req:
        payload += sch.ljust(16).encode() + i32(nt, nphi, jnu, one_d) + np.array(ds, np.float64).tobytes()
        payload += i32(len(mu)) + np.array(mu, np.float64).tobytes()
    out = run("quad", payload, EXE_RT)
    p, nray, vecs, wts = 0, [], [], []
    for _ in req:
        n = int(np.frombuffer(out[p:p + 4], np.int32)[0]); p += 4
        v = np.frombuffer(out[p:p + 24 * n], np.float64).reshape(n, 3); p += 24 * n
        w = np.frombuffer(out[p:p + 8 * n], np.float64); p += 8 * n
        nray.append(n)
        vecs.append(v)
        wts.append(w)
    mus = np.zeros((len(req), 5))
    for i, r in enumerate(req):
        mus[i, :len(r[6])] = r[6]
    save("rt_quadrature.npz", rt=True, in_scheme=np.array([r[0] for r in req]),
         in_int=np.array([r[1:5] for r in req], np.int32), in_ds=np.array([r[5] for r in req]),
         in_nmu=np.array([len(r[6]) for r in req], np.int32), in_mu=mus, nray=np.array(nray, np.int32),
         vec=np.concatenate(vecs), wts=np.concatenate(wts))


def gen_pchip(rng):
    n = 4000
    y = (10 ** rng.uniform(-12, -2, (n, 3))).astype(np.float32)
    y[:500] = np.sort(y[:500], axis=1)                  # rising with depth index
    y[500:1000] = np.sort(y[500:1000], axis=1)[:, ::-1]  # falling
    y[1000:1100, 1] = y[1000:1100, 0]                   # flat first interval
    y[1100:1200, 2] = y[1100:1200, 1]                   # flat second interval
    y[1200:1300] = y[1200:1300, :1]                     # constant
    y[1300:1400] = 0.0
    y[1400:1500] *= rng.choice([-1.0, 1.0], (100, 3)).astype(np.float32)   # negative (masers)
    ds3 = 10e5
    out = np.frombuffer(run("pchip", i32(n) + np.float64(ds3).tobytes() + y.tobytes(), EXE_RT), np.float32)
    save("rt_pchip_ghost.npz", rt=True, in_y=y, in_ds3=np.float64(ds3), ynew=out)


def gen_optab(rng):
    nnu, nt, ne_ = 40, 12, 9
    lam = np.sort(rng.uniform(3000.0, 9000.0, nnu))
    ang = 0.3
    rot = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    mu = np.array([8.6, 30.5])
    c0 = np.array([-1.1, -3.2])
    dc = np.array([0.2, 0.75])
    ti = np.arange(nt)[None, :, None]
    ei = np.arange(ne_)[None, None, :]
    li = np.log(lam)[:, None, None]
    val = (np.exp(0.3 * ti - 0.2 * ei + 0.5 * li - 4.0) * (1 + 0.3 * rng.random((nnu, nt, ne_))))
    scat = 0.05 * val * (1 + rng.random((nnu, nt, ne_)))
    val, scat = val.astype(np.float32), scat.astype(np.float32)
    nc, nf = 600, 70
    # query points in PC coordinates over the table and a little beyond, back to (T, ne)
    p1 = c0[0] + rng.uniform(-0.3, nt - 0.7, nc) * dc[0]
    p2 = c0[1] + rng.uniform(-0.3, ne_ - 0.7, nc) * dc[1]
    d = rot @ np.vstack([p1, p2])
    temp = np.exp(d[0] + mu[0])
    ne = np.exp(d[1] + mu[1])
    chi500 = (10 ** rng.uniform(-9, -5, nc)).astype(np.float32)
    lam_q = np.concatenate([[lam[0] - 10.0, lam[0], lam[-1], lam[-1] + 5.0],
                            np.sort(rng.uniform(lam[0], lam[-1], nf - 4))])
    F = lambda a: np.asfortranarray(a).tobytes(order="F")
    payload = (i32(nnu, nt, ne_) + mu.tobytes() + F(rot) + c0.tobytes() + dc.tobytes() + lam.tobytes()
               + F(val) + F(scat) + i32(nc, nf) + temp.tobytes() + ne.tobytes() + chi500.tobytes()
               + lam_q.tobytes())
    out = np.frombuffer(run("optab", payload, EXE_RT), np.float32)
    n3 = nnu * nt * ne_
    sl = [out[i * n3:(i + 1) * n3].reshape(ne_, nt, nnu).transpose(2, 1, 0) for i in range(6)]
    o = 6 * n3
    back = out[o:o + nf * nc].reshape(nc, nf)
    sc = out[o + nf * nc:o + 2 * nf * nc].reshape(nc, nf)
    save("hop_optab_sample.npz", rt=True, in_lam=lam, in_val=val, in_scat=scat, in_mu=mu, in_rot=rot, in_c0=c0,
         in_dc=dc, in_temp=temp, in_ne=ne, in_chi500=chi500, in_lam_q=lam_q, dl=sl[0], dt=sl[1], de=sl[2],
         sl=sl[3], st=sl[4], se=sl[5], back=back, scat=sc)


def gen_save(rng):
    nf, nx, ny, nmu = 37, 9, 7, 3
    nu = np.sort(rng.uniform(3.0e14, 1.2e15, nf))
    wmu = np.array([0.0833333333, 0.3101752434, 0.2731580899])
    I = (10 ** rng.uniform(-12, -2, (nmu, ny, nx, nf))).astype(np.float32)   # C order = Fortran (nf,nx,ny,nmu)
    out = np.frombuffer(run("save", i32(nf, nx, ny, nmu) + nu.tobytes() + wmu.tobytes() + I.tobytes(), EXE_RT),
                        np.float32)
    n3 = I.size
    i3 = out[:n3].reshape(nmu, ny, nx, nf)
    ie = out[n3:n3 + nf * nmu].reshape(nmu, nf)
    flux = out[n3 + nf * nmu:]
    save("rt_save_intensity.npz", rt=True, in_nu=nu, in_wmu=wmu, in_I=I, i3=i3, ie=ie, flux=flux)


def main():
    """python make_golden.py [name ...]: every fixture, or only the named ones (the
    seeded draws are the same either way: each fixture has its own generator)."""
    for exe in (EXE, EXE_RT):
        if not os.path.exists(exe):
            sys.exit(f"{exe} missing: run oracle/build_ref.sh first")
    only = set(sys.argv[1:])
    ONLY.update(only)
    rng = np.random.default_rng(20261202)
    gen_expneg(rng)
    gen_qlg(rng)
    gen_lineq(rng)
    gen_opac(rng, "hop_opac_equi_lte.npz", flags=0)
    gen_opac(rng, "hop_opac_equi_nlte.npz", flags=1 | 2, np_=2)
    gen_opac(rng, "hop_opac_equi_scat.npz", flags=4, nl=30, nf=200, m=(3, 2, 2))
    for i, (name, kw) in enumerate([("hop_opac_gen_lte.npz", dict(flags=0)),
                                    ("hop_opac_gen_nlte.npz", dict(flags=1 | 2, np_=2)),
                                    ("hop_opac_gen_scat.npz", dict(flags=4, nl=30, nf=200, m=(3, 2, 2)))]):
        if not only or name in only:
            gen_opac(np.random.default_rng(20261210 + i), name, general=True, **kw)
    for i, (name, kw) in enumerate([("hop_opac_equi_reach.npz", dict(flags=8)),
                                    ("hop_opac_equi_reach_scat.npz", dict(flags=8 | 4, nl=30, nf=200, m=(3, 2, 2))),
                                    ("hop_opac_gen_reach.npz", dict(flags=8, general=True)),
                                    ("hop_opac_gen_reach_nlte.npz", dict(flags=8 | 1 | 2, np_=2, general=True))]):
        if not only or name in only:
            gen_opac(np.random.default_rng(20261230 + i), name, **kw)
    gen_quad()
    gen_pchip(np.random.default_rng(20261220))
    gen_optab(np.random.default_rng(20261221))
    gen_save(np.random.default_rng(20261222))


if __name__ == "__main__":
    main()
