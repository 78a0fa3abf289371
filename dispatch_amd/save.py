"""The reference's on-disk result format for the emergent spectrum, so that readers of
dispatch runs (tumult's io_binary) take GPU results unchanged.

patch_t%save5 (tasks/patch_mod.f90:4248-4264) writes one array per patch as a raw
Fortran-order real(4) stream ``<run_dir>/save/<label>_<k1><k2>.bin`` (the patch key,
two three-digit fields, get_meta_name :4286-4308) and appends a line to
``<label>_meta.txt`` (save_meta :4311-4347): the file name right-aligned in 32
characters, the element size in bits (I4) and the shape (I10 each), under the header
line ``* filename ... bytes``.  rt_patch_t%save_intensity
(SRC/rt_patch_mod.f90:4151-4195) saves per frequency window

* ``ie``   (nf, 1, 1, nmu, n_set): the emergent intensity averaged over the surface,
  2 h nu^5 / (c^3 1e8) I [erg/s/cm^2/A/sr];
* ``i3``   (nf, nx, ny, nmu, n_set): the same per surface cell (save_resolved);
* ``flux`` (nf, 1, 1, 1, n_set): 2 pi sum_mu wmu ie, wmu the output rays' weights
  that already carry the cosine.

i3 is formed on the device; the surface and ray sums run on the host in the
reference's order, so the files match the reference's bit for bit.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from .constants import CC, CM_TO_AA, HH, PI

# 2 h / (c^3 1e8): occupation units to erg/s/cm^2/A/sr per nu^5 (rt_patch_mod.f90:4154)
CONVERSION = 2 * HH / (CC ** 3 * CM_TO_AA)


def file_name(label: str, key) -> str:
    """get_meta_name: '<label>_<|k1|:03d><|k2|:03d>.bin', blanks in label as '_'."""
    return label.replace(" ", "_") + "_" + "".join(f"{abs(int(k)):03d}" for k in key) + ".bin"


def meta_line(name: str, arr: np.ndarray) -> str:
    """save_meta's record '(A32, I4, n(I10))'."""
    return f"{name:>32s}{8 * arr.dtype.itemsize:4d}" + "".join(f"{d:10d}" for d in arr.shape)


META_HEADER = "*" + " filename" + " " * 21 + "bytes"


def save5(run_dir: str, label: str, arr, key=(1, 1)) -> str:
    """One patch array in the reference's format; returns the .bin path.  arr: a
    (device or host) tensor or array indexed as the Fortran array is (a[f, x, y, mu,
    set]); the stream is written in Fortran order, first index fastest."""
    a = arr.detach().cpu().numpy() if isinstance(arr, torch.Tensor) else np.asarray(arr)
    a = np.asarray(a, dtype="<f4")
    folder = os.path.join(run_dir, "save")
    os.makedirs(folder, exist_ok=True)
    name = file_name(label, key)
    meta = os.path.join(folder, label.replace(" ", "_") + "_meta.txt")
    new = not os.path.exists(meta)
    with open(meta, "a") as m:
        if new:
            m.write(META_HEADER + "\n")
        m.write(meta_line(name, a) + "\n")
    path = os.path.join(folder, name)
    with open(path, "wb") as f:
        f.write(np.asfortranarray(a).tobytes(order="F"))
    return path


def emergent_products(exits, nu, wmu, resolved=False):
    """ie, flux (and i3) of save_intensity from the exit planes of the output rays.

    exits: (nmu, nf, ny, nx) emergent intensities in occupation units (FP32, device),
    nu: (nf) frequencies [Hz], wmu: (nmu) the output rays' flux weights (cosine
    included).  Returns Fortran-shaped arrays: ie (nf, 1, 1, nmu, 1), flux
    (nf, 1, 1, 1, 1), i3 (nf, nx, ny, nmu, 1) or None.

    The reference's arithmetic, step for step (SRC/rt_patch_mod.f90:4165-4192):
    i3 = conversion nu^5 I formed in real(8) and rounded once to real(4) (on the
    device); ie the real(4) sum over the surface cells, y outer and x inner, one
    cell after another, then divided by real(nx ny); flux the real(4) sum over the
    rays of real(2 pi wmu, 4) ie.  The two sums run on the host (they are a few
    thousand vector adds of nmu x nf values, and their order is what fixes the
    result's last bit)."""
    nmu, nf, ny, nx = exits.shape
    scale64 = CONVERSION * np.asarray(nu, dtype=np.float64) ** 5
    scale = torch.as_tensor(scale64, dtype=torch.float64, device=exits.device)
    i3 = (scale[None, :, None, None] * exits.double()).float()            # (mu, f, y, x) real(4)
    cells = i3.permute(2, 3, 0, 1).contiguous().cpu().numpy()             # (y, x, mu, f)
    ie = np.zeros((nmu, nf), np.float32)
    for y in range(ny):
        for x in range(nx):
            ie += cells[y, x]
    ie /= np.float32(nx * ny)
    flux = np.zeros(nf, np.float32)
    for m in range(nmu):
        flux += np.float32(2 * PI * float(wmu[m])) * ie[m]
    ie_f = np.ascontiguousarray(ie.T).reshape(nf, 1, 1, nmu, 1)
    i3_f = i3.permute(1, 3, 2, 0).reshape(nf, nx, ny, nmu, 1) if resolved else None
    return ie_f, flux.reshape(nf, 1, 1, 1, 1), i3_f


def save_intensity(run_dir, exits, nu, wmu, key=(1, 1), resolved=False):
    """save_intensity for one window: 'ie', 'flux' (and 'i3') under run_dir/save."""
    ie, flux, i3 = emergent_products(exits, nu, wmu, resolved)
    out = dict(ie=save5(run_dir, "ie", ie, key))
    if i3 is not None:
        out["i3"] = save5(run_dir, "i3", i3, key)
    out["flux"] = save5(run_dir, "flux", flux, key)
    return out
