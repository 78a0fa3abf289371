"""A Fortran host (tests/fortran/sweep_driver.f90) calling the GPU sweep through
the ISO_C_BINDING module dispatch_amd/fortran/m3drt_mod.f90, the way dispatch.x's
rt_patch_t would: reference-layout arrays, device copies by the library's helpers.
The driver checks the constant-source fixed point itself and prints PASS."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "sweep_driver.x")


def test_fortran_host_drives_the_sweep():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built: run __graft_entry__.build() (needs amdflang)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "PASS" in r.stdout
