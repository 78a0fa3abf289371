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
    # HOP/const_mod.f90:39,40,56: h = 6.626176e-27, c = 2.99792458e10, 1e8 A per cm
    conv = 2 * 6.626176e-27 / (2.99792458e10 ** 3 * 1e8) * nu ** 5
    e = exits.double().numpy()
    # i3[f, x, y, mu] = conv(f) I(mu, f, y, x)
    np.testing.assert_allclose(i3[(2, 7)][..., 0], np.einsum("f,mfyx->fxym", conv, e), rtol=2e-6)
    ref_ie = np.einsum("f,mf->fm", conv, e.mean(axis=(2, 3)))
    np.testing.assert_allclose(ie[(2, 7)][:, 0, 0, :, 0], ref_ie, rtol=2e-6)
    np.testing.assert_allclose(ie[(2, 8)][:, 0, 0, :, 0], 2 * ref_ie, rtol=2e-6)
    ref_flux = 2 * 3.14159265359 * (ref_ie * wmu[None, :]).sum(axis=1)
    np.testing.assert_allclose(flux[(2, 7)][:, 0, 0, 0, 0], ref_flux, rtol=2e-6)


def test_products_are_the_reference_build():
    """i3, ie and flux bit for bit against save_intensity's statements compiled with
    the reference's const_mod (tests/golden/rt_save_intensity.npz, from
    oracle/ref_harness/rt_golden.f90): the real(8) product rounded once, the
    sequential real(4) surface sum, the real(4) ray sum."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rt_save_intensity.npz"))
    I = g["in_I"]                                         # (mu, y, x, f)
    nmu, ny, nx, nf = I.shape
    exits = torch.from_numpy(np.ascontiguousarray(I.transpose(0, 3, 1, 2)))
    ie, flux, i3 = save.emergent_products(exits, g["in_nu"], g["in_wmu"], resolved=True)
    assert np.array_equal(i3.numpy()[..., 0], g["i3"].transpose(3, 2, 1, 0))
    assert np.array_equal(ie[:, 0, 0, :, 0], g["ie"].T)
    assert np.array_equal(flux.ravel(), g["flux"])
    # the same sum with x outer and y inner differs in the last bits: the order matters
    xy = np.zeros((nmu, nf), np.float32)
    for x in range(nx):
        for y in range(ny):
            xy += g["i3"][:, y, x]
    assert not np.array_equal(xy / np.float32(nx * ny), g["ie"])
