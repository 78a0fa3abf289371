"""The oracle of rotate_array (oracle/m3d_rotate.c) pinned by exact answers: a
vertical ray leaves the cubes unchanged; a whole-cell shift (ray_offset integer, no
tilt) is a periodic roll in the direction the reference's modulo(x + up - 1, nx) + 1
takes; every plane's sum is conserved; a field linear in x and y is shifted exactly
where the stencil does not wrap; the shift grows linearly with the distance from the
pivot plane.  Each catches a plausible mistake (sign of the shift, a swapped weight,
the height measured from the wrong end)."""
import numpy as np

from oracle import oracle as O

DS = (1.0, 1.0, 0.5)


def cubes(seed, nk=2, nz=4, ny=6, nx=7, nf=3):
    return np.random.default_rng(seed).random((nk, nz, ny, nx, nf)).astype(np.float32)


def test_vertical_ray_is_identity():
    x = cubes(1)
    np.testing.assert_array_equal(O.rotate_array(x, (0.0, 0.0, 1.0), DS, 2.0), x)


def test_whole_cell_offset_is_a_roll():
    x = cubes(2)
    out = O.rotate_array(x, (0.0, 0.0, 1.0), DS, 2.0, ray_offset=(2.0, -3.0))
    # out(x, y) = in(x + 2, y - 3), periodically
    np.testing.assert_array_equal(out, np.roll(x, shift=(3, -2), axis=(2, 3)))


def test_plane_sums_are_conserved():
    x = cubes(3, nz=5)
    out = O.rotate_array(x, (0.37, -0.21, 0.9), DS, 2.0, ray_offset=(0.25, 0.5))
    np.testing.assert_allclose(out.astype(np.float64).sum(axis=(2, 3)), x.astype(np.float64).sum(axis=(2, 3)),
                               rtol=2e-6)


def test_linear_field_is_shifted_exactly_and_the_shift_follows_the_height():
    nz, ny, nx = 4, 40, 40
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    plane = (3.0 + 0.25 * xx - 0.125 * yy).astype(np.float32)
    x = np.broadcast_to(plane[None, None, :, :, None], (1, nz, ny, nx, 1)).copy()
    vec = np.array([0.3, -0.2, 0.8])
    b2p = 2.0
    out = O.rotate_array(x, vec, DS, b2p)
    for z in range(1, nz + 1):
        dk = (b2p - z * DS[2]) * vec[:2] / (vec[2] * np.array(DS[:2]))
        want = plane + 0.25 * dk[0] - 0.125 * dk[1]
        inner = (slice(5, ny - 5), slice(5, nx - 5))
        np.testing.assert_allclose(out[0, z - 1, :, :, 0][inner], want[inner], rtol=1e-6)
