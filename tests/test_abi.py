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


MIRRORS = {  # ctypes mirror in dispatch_amd/m3drt.py -> header typedef
    "SweepDesc": "m3drt_sweep_desc", "SweepBufs": "m3drt_sweep_bufs", "Linelist": "m3drt_linelist",
    "OpacityArgs": "m3drt_opacity_args", "RateSolveArgs": "m3drt_rate_solve_args",
    "RateSolveResult": "m3drt_rate_solve_result", "BbRatesArgs": "m3drt_bb_rates_args",
    "BbOpacityArgs": "m3drt_bb_opacity_args", "BfRatesArgs": "m3drt_bf_rates_args",
    "UnitsArgs": "m3drt_units_args", "EtaArgs": "m3drt_eta_args", "BackgroundArgs": "m3drt_background_args",
    "HermiteFields": "m3drt_hermite_fields", "AdjointBufs": "m3drt_adjoint_bufs",
    "B1dBatch": "m3drt_b1d_batch", "B1dTile": "m3drt_b1d_tile", "WindowArgs": "m3drt_window_args",
    "RotateArgs": "m3drt_rotate_args", "LcGhost": "m3drt_lc_ghost",
    "QuadratureArgs": "m3drt_quadrature_args", "BackgroundLteArgs": "m3drt_background_lte_args",
    "BoundaryArgs": "m3drt_boundary_args", "OptabTable": "m3drt_optab_table", "OptabQuery": "m3drt_optab_query",
    "Isotope": "m3drt_isotope", "Species": "m3drt_species", "LinelistOpacityArgs": "m3drt_linelist_opacity_args",
}


def test_struct_layout_matches_header():
    """Every field of every ctypes mirror sits at the C compiler's offset (a
    misnamed field fails to compile), and the sizes agree."""
    from dispatch_amd import m3drt
    lines, want = [], []
    for py, c in MIRRORS.items():
        S = getattr(m3drt, py)
        lines.append(f'  printf("%zu\\n", sizeof({c}));')
        want.append(C.sizeof(S))
        for name, _ in S._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({c}, {name}));')
            want.append(getattr(S, name).offset)
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"m3drt.h\"\nint main(void) {\n" + \
          "\n".join(lines) + "\n  return 0;\n}\n"
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        cfile = os.path.join(d, "t.c")
        open(cfile, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), cfile, "-o", exe], check=True)
        got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    assert got == want


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


def _header_structs():
    """(typedef name, [field names]) of every struct of include/m3drt.h."""
    import re
    src = open(os.path.join(ROOT, "include", "m3drt.h")).read()
    out = []
    for body, name in re.findall(r"typedef struct \{(.*?)\} (m3drt_[a-z0-9_]+);", src, re.S):
        fields = []
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        for line in body.split("\n"):
            line = line.strip().rstrip(";").replace("const ", "")
            if not line:
                continue
            rest = re.match(r"(?:int32_t|int64_t|float|double|void|uint8_t|char|m3drt_[a-z0-9_]+)\s+(.*)", line).group(1)
            fields += [re.sub(r"\[\d+\]", "", v.strip().lstrip("*").strip()) for v in rest.split(",")]
        out.append((name, fields))
    return out


def test_fortran_module_layouts_match_header():
    """Every bind(C) type of dispatch_amd/fortran/m3drt_mod.f90 has the C compiler's
    size and field offsets (amdflang against gcc)."""
    import shutil
    import tempfile
    fc = shutil.which("amdflang") or shutil.which("flang")
    if fc is None:
        pytest.skip("no Fortran compiler")
    structs = _header_structs()
    clines, flines = [], []
    for i, (name, fields) in enumerate(structs):
        clines.append(f'  printf("%zu\\n", sizeof({name}));')
        clines += [f'  printf("%zu\\n", offsetof({name}, {f}));' for f in fields]
        flines.append(f"  print '(i0)', c_sizeof(v{i})")
        flines += [f"  print '(i0)', transfer(c_loc(v{i}%{f}), 0_c_intptr_t) - transfer(c_loc(v{i}), 0_c_intptr_t)"
                   for f in fields]
    decl = "\n".join(f"  type({name}), target :: v{i}" for i, (name, _) in enumerate(structs))
    fsrc = "program layouts\n  use iso_c_binding\n  use m3drt\n  implicit none\n" + decl + "\n" + \
           "\n".join(flines) + "\nend program layouts\n"
    csrc = "#include <stdio.h>\n#include <stddef.h>\n#include \"m3drt.h\"\nint main(void) {\n" + \
           "\n".join(clines) + "\n  return 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "t.c"), "w").write(csrc)
        open(os.path.join(d, "t.f90"), "w").write(fsrc)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "t.c", "-o", "tc"], cwd=d, check=True)
        subprocess.run([fc, "-c", os.path.join(ROOT, "dispatch_amd", "fortran", "m3drt_mod.f90")], cwd=d, check=True)
        # the module procedures name library symbols this program never calls
        subprocess.run([fc, "t.f90", "m3drt_mod.o", "-Wl,--unresolved-symbols=ignore-all", "-o", "tf"], cwd=d,
                       check=True)
        want = subprocess.run([os.path.join(d, "tc")], capture_output=True, text=True, check=True).stdout.split()
        got = subprocess.run([os.path.join(d, "tf")], capture_output=True, text=True, check=True).stdout.split()
    assert got == want
