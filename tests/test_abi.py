"""The C ABI: every entry point include/m3drt.h declares is exported by
libm3drt.so, the ctypes structures match the header's layout, and the error
conventions of the reference's C ABI (HOP/standalone/hop_c_api.f90:1-18; int
status, message by *_last_error, no process exit) hold without a GPU.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "m3drt.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(m3drt_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    from dispatch_amd import m3drt
    return m3drt.lib()


def test_header_symbols_are_exported(lib):
    from dispatch_amd import m3drt
    out = subprocess.run(["nm", "-D", "--defined-only", m3drt.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (m3drt_\w+)", out))
    names = declared()
    assert len(names) >= 12
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)


def test_struct_layout_matches_header():
    # the C compiler's offsets of the descriptor fields vs the ctypes mirror
    from dispatch_amd import m3drt
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "m3drt.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(m3drt_sweep_desc), offsetof(m3drt_sweep_desc, ds),
         offsetof(m3drt_sweep_desc, wj), offsetof(m3drt_sweep_desc, cs_f),
         offsetof(m3drt_sweep_desc, eta_ld), sizeof(m3drt_sweep_bufs));
  return 0;
}'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    D = m3drt.SweepDesc
    assert vals == [C.sizeof(D), D.ds.offset, D.wj.offset, D.cs_f.offset, D.eta_ld.offset,
                    C.sizeof(m3drt.SweepBufs)]


def test_errors_are_status_codes(lib):
    from dispatch_amd import m3drt
    buf = C.create_string_buffer(256)
    # NULL context / descriptor: EINVAL with a message, no exit
    st = lib.m3drt_sc_sweep(None, None, None)
    assert st == 1
    assert lib.m3drt_last_error(buf, 256) == 0
    assert b"NULL" in buf.value
    assert lib.m3drt_create(None, 0, 0) == 1
    # a bad device number fails with a status as well (no GPU in the build container)
    h = C.c_void_p()
    st = lib.m3drt_create(C.byref(h), 9999, 0)
    assert st != 0 and not h.value
    assert lib.m3drt_last_error(buf, 256) == 0 and len(buf.value) > 0
    # strides helpers are pure host code
    d = m3drt.SweepDesc()
    d.nf, d.nx, d.ny, d.nz = 7, 5, 3, 2
    assert lib.m3drt_native_strides(C.byref(d)) == 0
    assert (d.cs_x, d.cs_y, d.cs_f, d.cs_z) == (1, 5, 15, 105)
    assert lib.m3drt_reference_strides(C.byref(d)) == 0
    assert (d.cs_f, d.cs_x, d.cs_y, d.cs_z) == (1, 7, 35, 105)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from dispatch_amd import m3drt
    monkeypatch.setattr(m3drt, "_lib", None)
    monkeypatch.setattr(m3drt, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(m3drt.M3drtError):
        m3drt.lib()
