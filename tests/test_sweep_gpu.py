"""GPU parity of the short-characteristics sweep (m3drt_sc_sweep) against the
oracle's restatement of short_char2 / long_char2 on the same seeded inputs.

FP32: the kernel forms the step weights with the hardware exponential and
reciprocal, the oracle with the reference's exp_neg and IEEE division, so they
agree to rounding, not bit for bit.  Tolerance: 1e-5 relative (the accuracy the
FP32 sweep must hold against FP64), with a floor of 1e-6 of the field's maximum
for values that cancel to near zero.  FP64: the kernel against the oracle's FP64
twin to 1e-11.
"""
import numpy as np
import pytest
import torch

from sweep_cases import interior, make_case, max_rel, plane_to_native, run_oracle, to_native

pytestmark = pytest.mark.gpu
RTOL32 = 1e-5


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def run_gpu(ctx, case, layout="native", dtype=torch.float32, split_rates=False, wj=0.25, full=True):
    gz = case["gz"]
    chi, S, I = case["chi"], case["S"], case["I"]
    if layout == "native":
        cub = lambda a: dev(to_native(interior(a)), dtype)
        pla = lambda a: dev(plane_to_native(a[gz]), dtype)
    else:
        cub = lambda a: dev(interior(a), dtype)
        pla = lambda a: dev(a[gz], dtype)
    c, s = cub(chi), cub(S)
    J = torch.full_like(c, 0.125)
    PsiJ = torch.full_like(c, 0.5)
    Iout = torch.zeros_like(c) if full else None
    Psi = torch.zeros_like(c) if full else None
    Irem = torch.zeros_like(c) if split_rates else None
    esc = torch.zeros_like(c) if split_rates else None
    exitI = torch.zeros_like(pla(chi))
    er = dev(case["eta_row"]) if "eta_row" in case else None
    em = dev(case["eta_mask"]) if "eta_mask" in case else None
    ctx.sc_sweep(c, s, pla(chi), pla(S), pla(I), dev(case["chi500"]), dev(case["chi500_up"]),
                 dev(case["g500"]), dir=case["dir"], ds=case["ds"], vec=case["vec"], layout=layout,
                 wj=wj, J=J, PsiJ=PsiJ, I_out=Iout, Psi_out=Psi, Irem_out=Irem, escape_out=esc,
                 exit_I=exitI, eta_row=er, eta_mask=em, split_rates=split_rates, get_rates=True)
    torch.cuda.synchronize()
    back = {}
    for k, v in dict(I=Iout, Psi=Psi, J=J, PsiJ=PsiJ, Irem=Irem, esc=esc).items():
        if v is None:
            continue
        a = v.double().cpu().numpy()
        back[k] = np.moveaxis(a, 1, 3) if layout == "native" else a
    e = exitI.double().cpu().numpy()
    back["exit"] = np.moveaxis(e, 0, 2) if layout == "native" else e
    return back


def check32(got, ref, what, rtol=RTOL32):
    ref = np.asarray(ref, np.float64)
    floor = 1e-6 * np.max(np.abs(ref)) if ref.size else 0.0
    err = max_rel(got, ref, floor)
    assert err < rtol, f"{what}: max rel err {err:.3e}"


CASES = [
    dict(seed=1, vec=(0.3, -0.4, 0.866)),
    dict(seed=2, vec=(-0.5, 0.2, -0.84)),
    dict(seed=3, vec=(0.61481017, 0.0, 0.78867513)),
    dict(seed=4, vec=(-0.48910, 0.84715, 0.21132487)),
    dict(seed=5, vec=(0.9, 0.2, 0.39), nx=17, ny=13, nz=7),
]


@pytest.mark.parametrize("kw", CASES)
@pytest.mark.parametrize("layout", ["native", "reference"])
def test_sweep_fp32_matches_oracle(ctx, kw, layout):
    kw = dict(kw)
    case = make_case(nf=kw.pop("nf", 16), nx=kw.pop("nx", 12), ny=kw.pop("ny", 10), nz=kw.pop("nz", 9), **kw)
    print(f"case seed={case['dir']} {kw} layout={layout}", flush=True)
    I, psi, _, _ = run_oracle(case)
    g = run_gpu(ctx, case, layout)
    check32(g["I"], interior(I), "I")
    check32(g["Psi"], psi, "Psi")
    # fused angle integration: J += wj*I, PsiJ += wj*Psi
    check32(g["J"], 0.125 + 0.25 * interior(I).astype(np.float64), "J")
    check32(g["PsiJ"], 0.5 + 0.25 * psi.astype(np.float64), "PsiJ")
    last = case["nz"] if case["dir"] > 0 else 1
    check32(g["exit"], I[last], "exit plane")


@pytest.mark.parametrize("seed,vec,maser", [(7, (0.2, 0.9, 0.39), False), (8, (0.1, -0.97, -0.211), True)])
def test_sweep_eta_and_split_rates(ctx, seed, vec, maser):
    case = make_case(nf=16, nx=11, ny=9, nz=8, seed=seed, vec=vec, eta=True, maser=maser)
    print(f"eta case seed={seed} maser={maser}", flush=True)
    I, psi, irem, esc = run_oracle(case, split_rates=True)
    g = run_gpu(ctx, case, split_rates=True)
    check32(g["I"], interior(I), "I")
    check32(g["Psi"], psi, "Psi")
    check32(g["Irem"], irem, "Irem")
    # escape of a flagged cell adds 0.5*Psi*(x_up - x)*rxb, a difference of two
    # nearly equal interpolated ratios: in FP32 it carries the cancellation of that
    # difference (the oracle's own FP32 escape is 2e-5..5e-5 off its FP64 one on
    # these cases).  Bar: the kernel is no further from FP64 than twice the
    # reference arithmetic in FP32, plus the usual 1e-5.
    esc64 = run_oracle(case, split_rates=True, fp64=True)[3]
    floor = 1e-6 * np.max(np.abs(esc64))
    own = max_rel(esc, esc64, floor)
    assert max_rel(g["esc"], esc64, floor) < RTOL32 + 2 * own, "escape"


@pytest.mark.parametrize("kw", CASES[:3])
def test_sweep_fp64_matches_oracle_fp64(ctx, kw):
    kw = dict(kw)
    case = make_case(nf=8, nx=12, ny=10, nz=9, **kw)
    I, psi, _, _ = run_oracle(case, fp64=True)
    g = run_gpu(ctx, case, dtype=torch.float64)
    assert max_rel(g["I"], interior(I), 1e-12) < 1e-11
    assert max_rel(g["Psi"], psi, 1e-12 * np.max(np.abs(psi))) < 1e-11


@pytest.mark.parametrize("nf", [1, 63, 200])
def test_sweep_1d_long_char2(ctx, nf):
    case = make_case(nf=nf, nx=1, ny=1, nz=40, seed=nf, vec=(0.0, 0.0, 0.5), eta=(nf == 63))
    print(f"1d nf={nf}", flush=True)
    from oracle import oracle as O
    I, psi, _, _ = O.long_char2(case["chi"], case["S"], case["I"], case["chi500"], case["g500"],
                                case["dir"], case["ds"], case["vec"], eta_row=case.get("eta_row"),
                                eta_mask=case.get("eta_mask"))
    for layout in ("native", "reference"):
        g = run_gpu(ctx, case, layout)
        check32(g["I"], interior(I), "I 1d")
        check32(g["Psi"], psi, "Psi 1d")


def test_sweep_plane_at_lds_limit(ctx):
    """The config-1 plane (120 x 120) on a few planes and frequencies: I and x sit
    in LDS, S of the upwind plane is re-read from global memory."""
    case = make_case(nf=4, nx=120, ny=120, nz=5, seed=12, vec=(-0.48910, 0.84715, 0.21132487))
    I, psi, _, _ = run_oracle(case)
    g = run_gpu(ctx, case)
    check32(g["I"], interior(I), "I 120x120")
    check32(g["Psi"], psi, "Psi 120x120")


def test_sweep_rejects_bad_arguments(ctx):
    from dispatch_amd import m3drt
    case = make_case(nf=2, nx=4, ny=4, nz=3, seed=0)
    c = dev(to_native(interior(case["chi"])))
    p = dev(plane_to_native(case["chi"][0]))
    with pytest.raises(m3drt.M3drtError, match="dir must follow"):
        ctx.sc_sweep(c, c, p, p, p, dev(case["chi500"]), dev(case["chi500_up"]), dev(case["g500"]),
                     dir=-1, ds=case["ds"], vec=case["vec"])
    with pytest.raises(m3drt.M3drtError, match="upwind"):
        ctx.sc_sweep(c, c, p, p, p, dev(case["chi500"]), None, dev(case["g500"]),
                     dir=1, ds=case["ds"], vec=case["vec"])


@pytest.mark.parametrize("nx,ny,fp64,psij,layout", [(60, 60, False, False, "native"), (60, 60, False, True, "native"),
                                                     (120, 120, False, False, "native"),
                                                     (96, 96, False, False, "native"), (96, 96, False, True, "native"),
                                                     (60, 60, False, False, "reference"),
                                                     (120, 120, False, False, "reference"),
                                                     (60, 60, True, False, "native"), (120, 120, True, True, "native")])
def test_sweep_exact_strips(ctx, nx, ny, fp64, psij, layout):
    """The straight-line kernel forms (only J, or J and PsiJ, asked for; the strip
    length divides ny): two columns per lane (native layout, J only), one column
    per lane (PsiJ, or the reference layout), the general kernel's exact strips
    (FP64): J, PsiJ and the exit plane against the oracle."""
    case = make_case(nf=3, nx=nx, ny=ny, nz=4, seed=nx + ny + int(fp64), vec=(0.61481017, 0.0, 0.78867513))
    print(f"exact strips {nx}x{ny} fp64={fp64} psij={psij} {layout}", flush=True)
    check_exact_strips(ctx, case, fp64, psij, layout)


@pytest.mark.parametrize("vec", [(0.31, 0.42, 0.85), (-0.31, 0.42, 0.85), (0.31, -0.42, 0.85), (-0.31, -0.42, 0.85),
                                 (0.25, -0.35, -0.90)])
def test_sweep_s_rows_in_lds(ctx, vec, monkeypatch, capfd):
    """120 x 120 planes with J and the exit plane as the only outputs: the form that
    holds the upwind S of the first strips' rows in LDS (beside the (I, x) plane)
    and reads the other strips' S from the cache.  Every sign of the horizontal
    step (both row offsets of the taps, so both placements of the periodic row),
    upward and downward, over 9 planes so the held rows are rewritten several
    times: J and the exit plane against the oracle."""
    case = make_case(nf=2, nx=120, ny=120, nz=9, seed=77, vec=vec)
    print(f"S rows in LDS 120x120 vec={vec}", flush=True)
    monkeypatch.setenv("M3DRT_VERBOSE", "1")
    check_exact_strips(ctx, case, False, False, "native")
    assert "LDS mode 4" in capfd.readouterr().err


@pytest.mark.parametrize("nx,ny,vec,maser,layout,strip", [(60, 60, (0.3, -0.4, 0.866), False, "native", 15),
                                                           (60, 60, (-0.25, 0.3, -0.92), True, "native", None),
                                                           (60, 60, (0.3, -0.4, 0.866), True, "native", 4),
                                                           (120, 120, (0.3, 0.2, 0.93), False, "reference", None)])
def test_sweep_split_outputs_fast_form(ctx, nx, ny, vec, maser, layout, strip, monkeypatch, capfd):
    """The ALI's split set alone (I, Psi, Irem and escape with rates, no J): the
    straight-line kernel's split form (S in LDS on the 60 x 60 plane, from the cache on
    120 x 120), with masing cells (negative chi) on two cases, against the oracle's
    short_char2 with split rates: I, Psi, Irem, escape and the exit plane.  On 60 x 60
    the default strip is 15 rows (checked on the first case); the third case takes
    strips of 4."""
    case = make_case(nf=3, nx=nx, ny=ny, nz=6, seed=nx + int(maser), vec=vec, maser=maser)
    print(f"split outputs fast form {nx}x{ny} maser={maser} {layout} strip={strip}", flush=True)
    if strip == 4:
        monkeypatch.setenv("M3DRT_FAST_L", "4")
    I, psi, irem, esc = run_oracle(case, split_rates=True)
    gz = case["gz"]
    if layout == "native":
        cub = lambda a: dev(to_native(interior(a)))
        pla = lambda a: dev(plane_to_native(a[gz]))
        nat = lambda t: np.moveaxis(t.double().cpu().numpy(), 1, 3)
        pnat = lambda t: np.moveaxis(t.double().cpu().numpy(), 0, 2)
    else:
        cub = lambda a: dev(interior(a))
        pla = lambda a: dev(a[gz])
        nat = pnat = lambda t: t.double().cpu().numpy()
    c, s = cub(case["chi"]), cub(case["S"])
    outs = {k: torch.full_like(c, 7.0) for k in ("I_out", "Psi_out", "Irem_out", "escape_out")}
    exitI = torch.zeros_like(pla(case["chi"]))
    monkeypatch.setenv("M3DRT_VERBOSE", "1")
    ctx.sc_sweep(c, s, pla(case["chi"]), pla(case["S"]), pla(case["I"]), dev(case["chi500"]),
                 dev(case["chi500_up"]), dev(case["g500"]), dir=case["dir"], ds=case["ds"], vec=case["vec"],
                 split_rates=True, get_rates=True, exit_I=exitI, layout=layout, **outs)
    torch.cuda.synchronize()
    err = capfd.readouterr().err
    assert "split outputs" in err
    if strip:
        assert f"strip {strip}," in err
    check32(nat(outs["I_out"]), interior(I), "I")
    check32(nat(outs["Psi_out"]), psi, "Psi")
    check32(nat(outs["Irem_out"]), irem, "Irem")
    check32(nat(outs["escape_out"]), esc, "escape")
    last = case["nz"] if case["dir"] > 0 else 1
    check32(pnat(exitI), I[last], "exit plane")


@pytest.mark.parametrize("nx,ny,vec,layout", [(96, 96, (0.35, 0.3, 0.887), "native"),
                                              (120, 120, (-0.2, -0.3, -0.93), "reference")])
def test_sweep_intensity_cube_fast_form(ctx, nx, ny, vec, layout, monkeypatch, capfd):
    """The intensity cube alone (I_out, no J, no rates: what an adjoint sweep
    recomputes), in the straight-line kernel's I-only form, against the oracle; the
    exit plane with it."""
    case = make_case(nf=3, nx=nx, ny=ny, nz=5, seed=nx + 3, vec=vec)
    print(f"intensity cube fast form {nx}x{ny} {layout}", flush=True)
    I, _, _, _ = run_oracle(case)
    gz = case["gz"]
    if layout == "native":
        cub = lambda a: dev(to_native(interior(a)))
        pla = lambda a: dev(plane_to_native(a[gz]))
        nat = lambda t: np.moveaxis(t.double().cpu().numpy(), 1, 3)
        pnat = lambda t: np.moveaxis(t.double().cpu().numpy(), 0, 2)
    else:
        cub = lambda a: dev(interior(a))
        pla = lambda a: dev(a[gz])
        nat = pnat = lambda t: t.double().cpu().numpy()
    c, s = cub(case["chi"]), cub(case["S"])
    Iout = torch.full_like(c, 7.0)
    exitI = torch.zeros_like(pla(case["chi"]))
    monkeypatch.setenv("M3DRT_VERBOSE", "1")
    ctx.sc_sweep(c, s, pla(case["chi"]), pla(case["S"]), pla(case["I"]), dev(case["chi500"]),
                 dev(case["chi500_up"]), dev(case["g500"]), dir=case["dir"], ds=case["ds"], vec=case["vec"],
                 I_out=Iout, exit_I=exitI, layout=layout)
    torch.cuda.synchronize()
    assert "I out" in capfd.readouterr().err
    check32(nat(Iout), interior(I), "I")
    last = case["nz"] if case["dir"] > 0 else 1
    check32(pnat(exitI), I[last], "exit plane")


def check_exact_strips(ctx, case, fp64, psij, layout):
    I, psi, _, _ = run_oracle(case, fp64=fp64)
    dt = torch.float64 if fp64 else torch.float32
    gz = case["gz"]
    if layout == "native":
        cub = lambda a: dev(to_native(interior(a)), dt)
        pla = lambda a: dev(plane_to_native(a[gz]), dt)
    else:
        cub = lambda a: dev(interior(a), dt)
        pla = lambda a: dev(a[gz], dt)
    c, s = cub(case["chi"]), cub(case["S"])
    J = torch.full_like(c, 0.125)
    PsiJ = torch.full_like(c, 0.5) if psij else None
    exitI = torch.zeros_like(pla(case["chi"]))
    ctx.sc_sweep(c, s, pla(case["chi"]), pla(case["S"]), pla(case["I"]), dev(case["chi500"]),
                 dev(case["chi500_up"]), dev(case["g500"]), dir=case["dir"], ds=case["ds"], vec=case["vec"],
                 wj=0.25, J=J, PsiJ=PsiJ, exit_I=exitI, get_rates=psij, layout=layout)
    torch.cuda.synchronize()
    if layout == "native":
        nat = lambda t: np.moveaxis(t.double().cpu().numpy(), 1, 3)
        pnat = lambda t: np.moveaxis(t.double().cpu().numpy(), 0, 2)
    else:
        nat = pnat = lambda t: t.double().cpu().numpy()
    tol = 1e-11 if fp64 else RTOL32
    check32(nat(J), 0.125 + 0.25 * interior(I).astype(np.float64), "J", rtol=tol)
    if psij:
        check32(nat(PsiJ), 0.5 + 0.25 * psi.astype(np.float64), "PsiJ", rtol=tol)
    last = case["nz"] if case["dir"] > 0 else 1
    check32(pnat(exitI), I[last], "exit plane", rtol=tol)


@pytest.mark.parametrize("layout", ["native", "reference"])
@pytest.mark.parametrize("vec,nz", [((0.0, 0.0, 0.78867513), 30), ((0.0, 0.0, -0.21132487), 30), ((0.0, 0.0, 1.0), 2)])
def test_sweep_quadratic_columns(ctx, layout, vec, nz):
    """long_char_quadratic along every column (M3DRT_QUADRATIC): I, Psi (= Lambda*
    / chi), J and the exit plane against the oracle's quad_step restatement."""
    from oracle import oracle as O
    case = make_case(nf=24, nx=5, ny=4, nz=nz, seed=40 + nz, vec=vec)
    print(f"quadratic {vec} nz={nz} {layout}", flush=True)
    I, psi = O.long_char_quadratic(case["chi"], case["S"], case["I"], case["dir"], case["ds"], case["vec"])
    gz = case["gz"]
    if layout == "native":
        cub = lambda a: dev(to_native(interior(a)))
        pla = lambda a: dev(plane_to_native(a[gz]))
        back = lambda t: np.moveaxis(t.double().cpu().numpy(), 1, 3)
        pback = lambda t: np.moveaxis(t.double().cpu().numpy(), 0, 2)
    else:
        cub = lambda a: dev(interior(a))
        pla = lambda a: dev(a[gz])
        back = pback = lambda t: t.double().cpu().numpy()
    c = cub(case["chi"])
    J = torch.full_like(c, 0.125)
    Iout, Psi = torch.zeros_like(c), torch.zeros_like(c)
    exitI = torch.zeros_like(pla(case["chi"]))
    ctx.sc_sweep(c, cub(case["S"]), pla(case["chi"]), pla(case["S"]), pla(case["I"]), None, None, None,
                 dir=case["dir"], ds=case["ds"], vec=case["vec"], layout=layout, wj=0.25, J=J, I_out=Iout,
                 Psi_out=Psi, exit_I=exitI, solver="quadratic")
    torch.cuda.synchronize()
    check32(back(Iout), interior(I), "I quadratic")
    check32(back(Psi), psi, "Lambda*/chi quadratic")
    check32(back(J), 0.125 + 0.25 * interior(I).astype(np.float64), "J quadratic")
    last = case["nz"] if case["dir"] > 0 else 1
    check32(pback(exitI), I[last], "exit plane quadratic")


@pytest.mark.parametrize("kw", [dict(seed=50, vec=(0.3, -0.4, 0.866)), dict(seed=51, vec=(-0.5, 0.2, -0.84)),
                                dict(seed=52, vec=(0.9, 0.2, 0.39), nx=17, ny=13, nz=7, split=True)])
def test_sweep_short_char_linear(ctx, kw):
    """short_char (M3DRT_LINEAR): chi, S, I from the 4x4 stencil with the inner-min
    clamp, against the oracle's restatement; split parts when asked."""
    from oracle import oracle as O
    kw = dict(kw)
    split = kw.pop("split", False)
    case = make_case(nf=kw.pop("nf", 12), nx=kw.pop("nx", 12), ny=kw.pop("ny", 10), nz=kw.pop("nz", 9), **kw)
    print(f"short_char {kw} split={split}", flush=True)
    I, psi, irem, esc = O.short_char(case["chi"], case["S"], case["I"], case["dir"], case["ds"], case["vec"],
                                     split_rates=split)
    gz = case["gz"]
    cub = lambda a: dev(to_native(interior(a)))
    pla = lambda a: dev(plane_to_native(a[gz]))
    back = lambda t: np.moveaxis(t.double().cpu().numpy(), 1, 3)
    c = cub(case["chi"])
    J = torch.full_like(c, 0.125)
    Iout, Psi = torch.zeros_like(c), torch.zeros_like(c)
    Irem = torch.zeros_like(c) if split else None
    Esc = torch.zeros_like(c) if split else None
    ctx.sc_sweep(c, cub(case["S"]), pla(case["chi"]), pla(case["S"]), pla(case["I"]), None, None, None,
                 dir=case["dir"], ds=case["ds"], vec=case["vec"], wj=0.25, J=J, I_out=Iout, Psi_out=Psi,
                 Irem_out=Irem, escape_out=Esc, split_rates=split, solver="linear")
    torch.cuda.synchronize()
    check32(back(Iout), interior(I), "I short_char")
    check32(back(Psi), psi, "Psi short_char")
    check32(back(J), 0.125 + 0.25 * interior(I).astype(np.float64), "J short_char")
    if split:
        check32(back(Irem), irem, "Irem short_char")
        check32(back(Esc), esc, "escape short_char")


@pytest.mark.parametrize("kw", [dict(seed=60, vec=(0.3, -0.4, 0.866)), dict(seed=61, vec=(-0.61, 0.2, -0.5)),
                                dict(seed=62, vec=(0.9, 0.2, 0.39), nx=17, ny=13, nz=7, split=True)])
def test_sweep_hermite(ctx, kw):
    """short_char_hermite (m3drt_sc_sweep_hermite) against the oracle's restatement:
    the quadratic step from the second plane on, the linear step on the first."""
    from oracle import oracle as O
    from hermite_cases import hermite_inputs
    kw = dict(kw)
    split = kw.pop("split", False)
    case = make_case(nf=kw.pop("nf", 10), nx=kw.pop("nx", 12), ny=kw.pop("ny", 10), nz=kw.pop("nz", 9), **kw)
    h = hermite_inputs(case, seed=case["nz"])
    print(f"hermite {kw} split={split}", flush=True)
    I, psi, irem, esc = O.short_char_hermite(case["chi"], case["S"], case["I"], h["beta"], h["pe4"], h["temp"],
                                             h["dpe"], h["dlnt"], h["hknu"], case["dir"], case["ds"], case["vec"],
                                             split_rates=split)
    gz = case["gz"]
    cub = lambda a: dev(to_native(interior(a)))
    pla = lambda a: dev(plane_to_native(a[gz]))
    back = lambda t: np.moveaxis(t.double().cpu().numpy(), 1, 3)
    c = cub(case["chi"])
    J = torch.full_like(c, 0.125)
    Iout, Psi = torch.zeros_like(c), torch.zeros_like(c)
    Irem = torch.zeros_like(c) if split else None
    Esc = torch.zeros_like(c) if split else None
    fields = dict(beta=dev(to_native(h["beta"])), pe4=dev(h["pe4"]), temp=dev(h["temp"]),
                  dpe_dx=dev(h["dpe"][0]), dpe_dy=dev(h["dpe"][1]), dpe_dz=dev(h["dpe"][2]),
                  dlnt_dx=dev(h["dlnt"][0]), dlnt_dy=dev(h["dlnt"][1]), dlnt_dz=dev(h["dlnt"][2]), hknu=dev(h["hknu"]))
    ctx.sc_sweep(c, cub(case["S"]), pla(case["chi"]), pla(case["S"]), pla(case["I"]), None, None, None,
                 dir=case["dir"], ds=case["ds"], vec=case["vec"], wj=0.25, J=J, I_out=Iout, Psi_out=Psi,
                 Irem_out=Irem, escape_out=Esc, split_rates=split, get_rates=True, hermite=fields)
    torch.cuda.synchronize()
    check32(back(Iout), interior(I), "I hermite")
    check32(back(Psi), psi, "Psi hermite")
    check32(back(J), 0.125 + 0.25 * interior(I).astype(np.float64), "J hermite")
    if split:
        check32(back(Irem), irem, "Irem hermite")
        check32(back(Esc), esc, "escape hermite")
