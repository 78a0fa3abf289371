"""A Fortran host (tests/fortran/sweep_driver.f90) calling the GPU sweep through
the ISO_C_BINDING module dispatch_amd/fortran/m3drt_mod.f90, the way dispatch.x's
rt_patch_t would: reference-layout arrays, device copies by the library's helpers.
The driver checks the constant-source fixed point itself and prints PASS; on a
random source it writes its inputs and the J and exit plane the GPU returned, and
the oracle's short_char2 on those inputs must agree to 1e-5."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "sweep_driver.x")


def test_fortran_host_drives_the_sweep(tmp_path):
    import numpy as np
    from oracle import oracle as O
    from sweep_cases import max_rel
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built: run __graft_entry__.build() (needs amdflang)")
    out = str(tmp_path / "case2.bin")
    r = subprocess.run([EXE, out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "PASS" in r.stdout and "WROTE" in r.stdout
    raw = open(out, "rb").read()
    nf, nx, ny, nz = np.frombuffer(raw[:16], np.int32)
    ds, vec, wmu2 = np.frombuffer(raw[16:40], np.float64), np.frombuffer(raw[40:64], np.float64), \
        np.frombuffer(raw[64:72], np.float64)[0]
    f32 = np.frombuffer(raw[72:], np.float32)
    sizes = dict(chi=(nz, ny, nx, nf), S=(nz, ny, nx, nf), ghost=(3, ny, nx, nf), c500=(nz, ny, nx),
                 cup=(nz, ny, nx), g500=(ny, nx), J=(nz, ny, nx, nf), ex=(ny, nx, nf))
    a, p = {}, 0
    for k, shp in sizes.items():
        n = int(np.prod(shp))
        a[k] = f32[p:p + n].reshape(shp)      # Fortran (f, x, y, z) read as C (z, y, x, f)
        p += n
    assert p == f32.size
    z = np.zeros((1, ny, nx, nf), np.float32)
    work = lambda cube, g: np.ascontiguousarray(np.concatenate([g[None], cube, z]))
    I = O.short_char2(work(a["chi"], a["ghost"][0]), work(a["S"], a["ghost"][1]),
                      work(np.zeros_like(a["chi"]), a["ghost"][2]), a["c500"], a["cup"], a["g500"], 1, ds, vec,
                      get_rates=False)[0]
    Jref = (np.float32(wmu2) * I[1:-1]).astype(np.float64)
    assert max_rel(a["J"], Jref) < 1e-5
    assert max_rel(a["ex"], I[nz].astype(np.float64)) < 1e-5
    assert float(np.abs(a["J"]).max()) > 0
