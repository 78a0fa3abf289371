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

Arrays are formed on the device in FP32, in the reference's order of operations, and
copied to the host once.
"""
from __future__ import annotations

import os

import numpy as np
import torch

H_PLANCK = 6.62607015e-27
C_LIGHT = 2.99792458e10
CM_TO_AA = 1e8
PI = 3.14159265359          # the reference's pi (HOP/const_mod.f90)


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
    (nf, 1, 1, 1, 1), i3 (nf, nx, ny, nmu, 1) or None."""
    nmu, nf, ny, nx = exits.shape
    conv = torch.as_tensor(2 * H_PLANCK / (C_LIGHT ** 3 * CM_TO_AA), dtype=torch.float64)
    scale = (conv * torch.as_tensor(nu, dtype=torch.float64) ** 5).float().to(exits.device)
    i3 = exits.float() * scale[None, :, None, None]                  # (mu, f, y, x) real(4)
    ie = i3.sum(dim=(2, 3), dtype=torch.float32) / np.float32(nx * ny)   # (mu, f)
    w = torch.as_tensor([np.float32(2 * PI * float(v)) for v in wmu], dtype=torch.float32, device=exits.device)
    flux = torch.zeros(nf, dtype=torch.float32, device=exits.device)
    for m in range(nmu):
        flux = flux + w[m] * ie[m]
    ie_f = ie.t().reshape(nf, 1, 1, nmu, 1)
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
