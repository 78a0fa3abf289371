"""The emergent-spectrum files in the reference's save format (dispatch_amd.save):
the record layout of save_meta and get_meta_name, Fortran-order streams read back
the way tumult's io_binary.load_patch_files reads them (restated here: meta lines,
'<f{bits/8}' memmaps of the listed shape in order 'F', keys from the three-digit
fields of the name), and the ie / flux definitions of save_intensity."""
import os
import textwrap

import numpy as np
import torch

from dispatch_amd import save


def read_like_tumult(folder, kind):
    dct = {}
    with open(os.path.join(folder, f"{kind}_meta.txt")) as meta:
        for line in meta:
            if line.strip().startswith("*"):
                continue
            info = line.split()
            dims = tuple(map(int, info[2:]))
            key = tuple(map(int, textwrap.wrap(info[0].split(".")[0].split("_")[-1], 3)))
            dct[key] = np.memmap(os.path.join(folder, info[0]), dtype=f"<f{int(info[1]) // 8}", mode="r",
                                 shape=dims, order="F")
    return dct


def test_meta_records_follow_save_meta():
    a = np.zeros((1024, 1, 1, 5, 1), np.float32)
    assert save.file_name("ie", (3, 12)) == "ie_003012.bin"
    assert save.file_name("my label", (-1, 2)) == "my_label_001002.bin"
    line = save.meta_line("ie_003012.bin", a)
    assert line == " " * 19 + "ie_003012.bin" + "  32" + "      1024" + "         1" * 2 + "         5" + "         1"
    assert save.META_HEADER == "* filename                     bytes"


def test_intensity_files_read_back_as_tumult_reads_them(tmp_path):
    g = torch.Generator().manual_seed(5)
    nmu, nf, ny, nx = 3, 40, 6, 5
    exits = torch.rand((nmu, nf, ny, nx), generator=g, dtype=torch.float32) + 0.1
    nu = 2.99792458e10 / (5000e-8 + 1e-10 * np.arange(nf))
    wmu = np.array([0.2, 0.3, 0.5]) * np.array([0.3, 0.6, 0.9])
    run = str(tmp_path)
    save.save_intensity(run, exits, nu, wmu, key=(2, 7), resolved=True)
    save.save_intensity(run, exits * 2, nu, wmu, key=(2, 8))
    folder = os.path.join(run, "save")
    ie = read_like_tumult(folder, "ie")
    flux = read_like_tumult(folder, "flux")
    i3 = read_like_tumult(folder, "i3")
    assert set(ie) == {(2, 7), (2, 8)} and set(i3) == {(2, 7)}
    conv = 2 * save.H_PLANCK / (save.C_LIGHT ** 3 * save.CM_TO_AA) * nu ** 5
    e = exits.double().numpy()
    # i3[f, x, y, mu] = conv(f) I(mu, f, y, x)
    np.testing.assert_allclose(i3[(2, 7)][..., 0], np.einsum("f,mfyx->fxym", conv, e), rtol=2e-6)
    ref_ie = np.einsum("f,mf->fm", conv, e.mean(axis=(2, 3)))
    np.testing.assert_allclose(ie[(2, 7)][:, 0, 0, :, 0], ref_ie, rtol=2e-6)
    np.testing.assert_allclose(ie[(2, 8)][:, 0, 0, :, 0], 2 * ref_ie, rtol=2e-6)
    ref_flux = 2 * save.PI * (ref_ie * wmu[None, :]).sum(axis=1)
    np.testing.assert_allclose(flux[(2, 7)][:, 0, 0, 0, 0], ref_flux, rtol=2e-6)
