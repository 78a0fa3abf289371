"""Builds libm3drt.so (the HIP kernels and the C ABI) for gfx950, in-tree.

    python -m dispatch_amd.build          # or: __graft_entry__.build()

Every source under dispatch_amd/csrc is compiled by hipcc into one shared library
dispatch_amd/libm3drt.so.  The Fortran interface module is compiled with amdflang
when it is present (dispatch_amd/fortran/).
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libm3drt.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 (v_pk_*_f32) stays off: hipcc would otherwise form it from
    # float2-shaped code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    p = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libm3drt")
    return p


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def build(verbose: bool = False) -> str:
    """Compile every source (in parallel) and link libm3drt.so; returns its path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [hipcc()] + HIPCC_FLAGS + ["-I", INCLUDE, "-I", CSRC, "-c", src, "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        msgs = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
        if p.returncode != 0:
            failed.append(src)
            print("\n".join(msgs), file=sys.stderr)
        elif verbose and msgs:
            print("\n".join(msgs))
    if failed:
        raise RuntimeError("hipcc failed on: " + ", ".join(os.path.basename(f) for f in failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


def build_fortran(verbose: bool = False):
    """Compile the ISO_C_BINDING module with amdflang, when amdflang exists."""
    src = os.path.join(HERE, "fortran", "m3drt_mod.f90")
    fc = shutil.which("amdflang") or os.path.join(ROCM, "bin", "amdflang")
    if not (os.path.exists(src) and os.path.exists(fc)):
        return None
    outdir = os.path.join(HERE, "fortran", "_build")
    os.makedirs(outdir, exist_ok=True)
    obj = os.path.join(outdir, "m3drt_mod.o")
    cmd = [fc, "-O2", "-fPIC", "-module-dir", outdir, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    # the Fortran host driver the GPU tests run (tests/fortran)
    drv = os.path.join(ROOT, "tests", "fortran", "sweep_driver.f90")
    if os.path.exists(drv):
        exe_dir = os.path.join(ROOT, "tests", "fortran", "_build")
        os.makedirs(exe_dir, exist_ok=True)
        exe = os.path.join(exe_dir, "sweep_driver.x")
        cmd = [fc, "-O2", "-I", outdir, drv, obj, "-L", HERE, "-lm3drt",
               "-Wl,-rpath," + os.path.relpath(HERE, exe_dir).join(["$ORIGIN/", ""]), "-o", exe]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, cwd=exe_dir)
    return obj


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
