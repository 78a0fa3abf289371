"""GPU rotate_array (m3drt_rotate_array) against the oracle's restatement, bit for
bit, in both layouts; and the output pass it serves: chi and S rotated for an
inclined ray, then long_char_quadratic along the columns (M3DRT_QUADRATIC), against
the oracle's rotation and quad_step restatement."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from sweep_cases import make_case

pytestmark = pytest.mark.gpu
DS = (1.0, 1.0, 0.5)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.mark.parametrize("layout", ["reference", "native"])
@pytest.mark.parametrize("vec,off", [((0.37, -0.21, 0.9), (0.25, 0.5)), ((-0.6, 0.35, -0.72), (0.0, 0.0)),
                                     ((0.0, 0.0, 1.0), (1.0, -2.0))])
def test_rotate_matches_oracle(ctx, layout, vec, off):
    rng = np.random.default_rng(7)
    x = rng.random((2, 9, 10, 12, 5)).astype(np.float32)       # (k, z, y, x, f)
    print(f"rotate {layout} vec={vec} offset={off}", flush=True)
    ref = O.rotate_array(x, vec, DS, 3.0, off)
    if layout == "reference":
        xi = dev(x)
        back = lambda t: t.cpu().numpy()
    else:
        xi = dev(np.moveaxis(x, 4, 2))                           # (k, z, f, y, x)
        back = lambda t: np.moveaxis(t.cpu().numpy(), 2, 4)
    out = torch.full_like(xi, -1.0)
    ctx.rotate_array(xi, out, vec, DS, 3.0, off, layout=layout)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(back(out), ref)


@pytest.mark.parametrize("flat", [False, True])
def test_rotate_rows_wider_than_a_wave(ctx, flat, monkeypatch):
    """The native layout's row kernel on a plane of 130 columns (two passes of a
    wave's lanes along x, the wrap in the second) and 7 rows, three frequencies, two
    fields; and the flat-index kernel (M3DRT_ROTATE_FLAT) on the same case."""
    rng = np.random.default_rng(11)
    x = rng.random((2, 5, 7, 130, 3)).astype(np.float32)       # (k, z, y, x, f)
    vec, off = (0.52, -0.31, 0.8), (0.75, -0.25)
    print(f"rotate rows 130x7 flat={flat}", flush=True)
    if flat:
        monkeypatch.setenv("M3DRT_ROTATE_FLAT", "1")
    ref = O.rotate_array(x, vec, DS, 2.5, off)
    xi = dev(np.moveaxis(x, 4, 2))                               # (k, z, f, y, x)
    out = torch.full_like(xi, -1.0)
    ctx.rotate_array(xi, out, vec, DS, 2.5, off, layout="native")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(np.moveaxis(out.cpu().numpy(), 2, 4), ref)


def test_output_pass_on_rotated_columns(ctx):
    vec = (0.3, -0.25, 0.92)
    case = make_case(nf=6, nx=9, ny=8, nz=12, seed=61, vec=vec)
    v = case["vec"]
    nz = case["nz"]
    b2p = nz * DS[2]
    print("output pass: rotate chi, S; quadratic columns", flush=True)
    work = np.stack([case["chi"][1:-1], case["S"][1:-1]]).astype(np.float32)   # (2, nz, ny, nx, nf)
    rot = O.rotate_array(work, v, case["ds"], b2p)
    chi, S = case["chi"].copy(), case["S"].copy()
    chi[1:-1], S[1:-1] = rot[0], rot[1]
    I, _ = O.long_char_quadratic(chi, S, case["I"], case["dir"], case["ds"], v)
    w = dev(work)
    r = torch.empty_like(w)
    ctx.rotate_array(w, r, v, case["ds"], b2p)
    gz = case["gz"]
    exitI = torch.zeros_like(dev(case["chi"][gz]))
    ctx.sc_sweep(r[0], r[1], dev(case["chi"][gz]), dev(case["S"][gz]), dev(case["I"][gz]), None, None, None,
                 dir=case["dir"], ds=case["ds"], vec=v, layout="reference", exit_I=exitI, solver="quadratic")
    torch.cuda.synchronize()
    last = nz if case["dir"] > 0 else 1
    got, want = exitI.double().cpu().numpy(), I[last].astype(np.float64)
    assert np.max(np.abs(got - want) / np.abs(want)) < 1e-5
