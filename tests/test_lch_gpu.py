"""GPU parity of the cubic-Bezier column sweep (m3drt_lc_sweep_hermite, the
reference's long_char_hermite) against the oracle's restatement on the same seeded
inputs: I, Psi (the Lambda diagonal over chi), J, PsiJ and the exit plane, for the
three boundary kinds below the first plane, both directions, nz = 2 (the last
plane's parabola reaching the ghost) and longer columns, in the native and the
reference layout.

FP32: the kernel forms e^-dt with the hardware exponential, the oracle with the
reference's exp_neg, so they agree to rounding; tolerance 1e-5 relative with a
floor of 1e-6 of the field's maximum, as the other FP32 sweeps.
"""
import numpy as np
import pytest
import torch

from hermite_cases import hermite_inputs
from sweep_cases import interior, make_case, max_rel, plane_to_native, to_native

pytestmark = pytest.mark.gpu
RTOL32 = 1e-5


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def check32(got, ref, what, rtol=RTOL32):
    ref = np.asarray(ref, np.float64)
    floor = 1e-6 * np.max(np.abs(ref)) if ref.size else 0.0
    err = max_rel(got, ref, floor)
    assert err < rtol, f"{what}: max rel err {err:.3e}"


def ghost_thermo(h, case, kind):
    if kind == "top":
        return None
    first = 0 if case["dir"] > 0 else -1
    return dict(pe=h["pe4"][first] * np.float32(1.1), dpe_dz=h["dpe"][2][first],
                temp=h["temp"][first] - np.float32(40.0), dlnt_dz=h["dlnt"][2][first])


@pytest.mark.parametrize("layout", ["native", "reference"])
@pytest.mark.parametrize("kind", ["top", "bottom", "upstream"])
@pytest.mark.parametrize("vec,nz", [((0.0, 0.0, 0.78867513), 12), ((0.0, 0.0, -0.5), 12), ((0.0, 0.0, 1.0), 2),
                                    ((0.0, 0.0, -0.21132487), 2)])
def test_lc_sweep_hermite(ctx, layout, kind, vec, nz):
    from oracle import oracle as O
    case = make_case(nf=20, nx=5, ny=4, nz=nz, seed=70 + nz, vec=vec, top_ghost_zero=(kind == "top"))
    h = hermite_inputs(case, seed=nz + 1)
    g = ghost_thermo(h, case, kind)
    print(f"long_char_hermite {kind} {vec} nz={nz} {layout}", flush=True)
    og = {} if g is None else dict(gpe=g["pe"], gdpe=g["dpe_dz"], gtemp=g["temp"], gdlnt=g["dlnt_dz"])
    I, psi = O.long_char_hermite(case["chi"], case["S"], case["I"], h["beta"], h["pe4"], h["temp"], h["dpe"][2],
                                 h["dlnt"][2], h["hknu"], case["dir"], case["ds"], case["vec"], ghost=kind, **og)
    assert np.all(np.isfinite(I)) and np.all(np.isfinite(psi))
    gz = case["gz"]
    if layout == "native":
        cub = lambda a: dev(to_native(a))
        pla = lambda a: dev(plane_to_native(a[gz]))
        back = lambda t: np.moveaxis(t.double().cpu().numpy(), 1, 3)
        pback = lambda t: np.moveaxis(t.double().cpu().numpy(), 0, 2)
    else:
        cub = lambda a: dev(a)
        pla = lambda a: dev(a[gz])
        back = pback = lambda t: t.double().cpu().numpy()
    c = cub(interior(case["chi"]))
    J = torch.full_like(c, 0.125)
    PsiJ = torch.full_like(c, 0.5)
    Iout, Psi = torch.zeros_like(c), torch.zeros_like(c)
    exitI = torch.zeros_like(pla(case["chi"]))
    fields = dict(beta=cub(h["beta"]), pe4=dev(h["pe4"]), temp=dev(h["temp"]), dpe_dz=dev(h["dpe"][2]),
                  dlnt_dz=dev(h["dlnt"][2]), hknu=dev(h["hknu"]))
    gf = None if g is None else {k: dev(v) for k, v in g.items()}
    ctx.lc_sweep_hermite(c, cub(interior(case["S"])), pla(case["chi"]), pla(case["S"]), pla(case["I"]),
                         dir=case["dir"], ds=case["ds"], vec=case["vec"], fields=fields, ghost=kind, ghost_fields=gf,
                         layout=layout, wj=0.25, J=J, PsiJ=PsiJ, I_out=Iout, Psi_out=Psi, exit_I=exitI)
    torch.cuda.synchronize()
    check32(back(Iout), interior(I), "I")
    check32(back(Psi), psi, "Psi")
    check32(back(J), 0.125 + 0.25 * interior(I).astype(np.float64), "J")
    check32(back(PsiJ), 0.5 + 0.25 * psi.astype(np.float64), "PsiJ")
    last = case["nz"] if case["dir"] > 0 else 1
    check32(pback(exitI), I[last], "exit plane")


def test_lc_sweep_hermite_rejects_bad_input(ctx):
    """Errors are status codes: nz < 2, a missing ghost state, split rates."""
    from dispatch_amd.m3drt import M3drtError
    case = make_case(nf=4, nx=3, ny=2, nz=1, seed=1, vec=(0.0, 0.0, 1.0))
    cub = lambda a: dev(to_native(a))
    pla = lambda a: dev(plane_to_native(a[case["gz"]]))
    c = cub(interior(case["chi"]))
    one = torch.ones(1, 2, 3, dtype=torch.float32, device="cuda")
    fields = dict(beta=torch.full_like(c, 0.1), pe4=one, temp=one * 5000, dpe_dz=one * 0, dlnt_dz=one * 0,
                  hknu=torch.full((4,), 1e4, dtype=torch.float32, device="cuda"))
    with pytest.raises(M3drtError, match="nz >= 2"):
        ctx.lc_sweep_hermite(c, cub(interior(case["S"])), pla(case["chi"]), pla(case["S"]), pla(case["I"]), dir=1,
                             ds=case["ds"], vec=case["vec"], fields=fields, I_out=torch.zeros_like(c))
    with pytest.raises(M3drtError, match="ghost_fields"):
        ctx.lc_sweep_hermite(c, cub(interior(case["S"])), pla(case["chi"]), pla(case["S"]), pla(case["I"]), dir=1,
                             ds=case["ds"], vec=case["vec"], fields=fields, ghost="bottom")


@pytest.mark.parametrize("layout", ["native", "reference"])
@pytest.mark.parametrize("vec,nz,split", [((0.0, 0.0, 0.78867513), 12, True), ((0.0, 0.0, -0.5), 9, False),
                                          ((0.0, 0.0, -1.0), 1, True)])
def test_long_char_linear_columns(ctx, layout, vec, nz, split):
    """long_char_linear along every column (M3DRT_LONG_LINEAR), masers included: I,
    Psi, the split parts, J, PsiJ and the exit plane against the oracle."""
    from oracle import oracle as O
    case = make_case(nf=24, nx=5, ny=4, nz=nz, seed=90 + nz, vec=vec, maser=True)
    print(f"long_char_linear {vec} nz={nz} split={split} {layout}", flush=True)
    I, psi, irem, esc = O.long_char_linear(case["chi"], case["S"], case["I"], case["dir"], case["ds"], case["vec"],
                                           split_rates=split)
    gz = case["gz"]
    if layout == "native":
        cub = lambda a: dev(to_native(a))
        pla = lambda a: dev(plane_to_native(a[gz]))
        back = lambda t: np.moveaxis(t.double().cpu().numpy(), 1, 3)
        pback = lambda t: np.moveaxis(t.double().cpu().numpy(), 0, 2)
    else:
        cub = lambda a: dev(a)
        pla = lambda a: dev(a[gz])
        back = pback = lambda t: t.double().cpu().numpy()
    c = cub(interior(case["chi"]))
    J, PsiJ = torch.full_like(c, 0.125), torch.full_like(c, 0.5)
    Iout, Psi = torch.zeros_like(c), torch.zeros_like(c)
    Irem = torch.zeros_like(c) if split else None
    Esc = torch.zeros_like(c) if split else None
    exitI = torch.zeros_like(pla(case["chi"]))
    ctx.sc_sweep(c, cub(interior(case["S"])), pla(case["chi"]), pla(case["S"]), pla(case["I"]), None, None, None,
                 dir=case["dir"], ds=case["ds"], vec=case["vec"], layout=layout, wj=0.25, J=J, PsiJ=PsiJ,
                 I_out=Iout, Psi_out=Psi, Irem_out=Irem, escape_out=Esc, exit_I=exitI, split_rates=split,
                 solver="long_linear")
    torch.cuda.synchronize()
    check32(back(Iout), interior(I), "I")
    check32(back(Psi), psi, "Psi")
    check32(back(J), 0.125 + 0.25 * interior(I).astype(np.float64), "J")
    check32(back(PsiJ), 0.5 + 0.25 * psi.astype(np.float64), "PsiJ")
    if split:
        check32(back(Irem), irem, "Irem")
        check32(back(Esc), esc, "escape")
    last = case["nz"] if case["dir"] > 0 else 1
    check32(pback(exitI), I[last], "exit plane")


@pytest.mark.parametrize("layout", ["native", "reference"])
@pytest.mark.parametrize("kind", ["top", "bottom", "upstream"])
@pytest.mark.parametrize("vec,nz", [((0.0, 0.0, 0.78867513), 12), ((0.0, 0.0, -0.5), 3), ((0.0, 0.0, 1.0), 2),
                                    ((0.0, 0.0, -0.21132487), 2)])
def test_lc_sweep_hermite2(ctx, layout, kind, vec, nz):
    """long_char_hermite2 (m3drt_lc_sweep_hermite2) against the oracle: I, Psi, J,
    PsiJ, the exit plane; masers included (the FP64 step floor and the clamp)."""
    from oracle import oracle as O
    case = make_case(nf=20, nx=5, ny=4, nz=nz, seed=80 + nz, vec=vec, top_ghost_zero=(kind == "top"),
                     maser=(nz == 12))
    h = hermite_inputs(case, seed=nz + 2)
    g = ghost_thermo(h, case, kind)
    print(f"long_char_hermite2 {kind} {vec} nz={nz} {layout}", flush=True)
    og = {} if g is None else dict(gpe=g["pe"], gdpe=g["dpe_dz"])
    I, psi = O.long_char_hermite2(case["chi"], case["S"], case["I"], h["pe4"], h["dpe"][2], case["dir"], case["ds"],
                                  case["vec"], ghost=kind, **og)
    assert np.all(np.isfinite(I)) and np.all(np.isfinite(psi))
    gz = case["gz"]
    if layout == "native":
        cub = lambda a: dev(to_native(a))
        pla = lambda a: dev(plane_to_native(a[gz]))
        back = lambda t: np.moveaxis(t.double().cpu().numpy(), 1, 3)
        pback = lambda t: np.moveaxis(t.double().cpu().numpy(), 0, 2)
    else:
        cub = lambda a: dev(a)
        pla = lambda a: dev(a[gz])
        back = pback = lambda t: t.double().cpu().numpy()
    c = cub(interior(case["chi"]))
    J, PsiJ = torch.full_like(c, 0.125), torch.full_like(c, 0.5)
    Iout, Psi = torch.zeros_like(c), torch.zeros_like(c)
    exitI = torch.zeros_like(pla(case["chi"]))
    fields = dict(pe4=dev(h["pe4"]), dpe_dz=dev(h["dpe"][2]))
    gf = None if g is None else dict(pe=dev(g["pe"]), dpe_dz=dev(g["dpe_dz"]))
    ctx.lc_sweep_hermite(c, cub(interior(case["S"])), pla(case["chi"]), pla(case["S"]), pla(case["I"]),
                         dir=case["dir"], ds=case["ds"], vec=case["vec"], fields=fields, ghost=kind, ghost_fields=gf,
                         layout=layout, wj=0.25, J=J, PsiJ=PsiJ, I_out=Iout, Psi_out=Psi, exit_I=exitI, hybrid=True)
    torch.cuda.synchronize()
    check32(back(Iout), interior(I), "I")
    check32(back(Psi), psi, "Psi")
    check32(back(J), 0.125 + 0.25 * interior(I).astype(np.float64), "J")
    check32(back(PsiJ), 0.5 + 0.25 * psi.astype(np.float64), "PsiJ")
    last = case["nz"] if case["dir"] > 0 else 1
    check32(pback(exitI), I[last], "exit plane")
