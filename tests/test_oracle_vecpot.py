"""The vector-potential curl that builds every initial field
(pluto_amd/problems.py:curl_edges) against PLUTO's own VectorPotentialDiff
(vec_pot_diff.c:139-171), driven by Startup() (startup.c:121-131,177-187)
in the reference built with ASSIGN_VECTOR_POTENTIAL YES.

* golden: tests/golden/vecpot_*.npz hold a random edge potential and the
  staggered and cell-centred field the reference made from it (3D on a
  non-cubic 8x7x6 box with unequal cell widths; 2D with only A_z);
* live, where oracle/_ref is built: another random potential, run now;
* a curl with a plausible mistake (a neighbour taken on the wrong side, a
  dropped term) differs from the reference, so these checks have teeth.
"""
import numpy as np
import pytest

from conftest import golden
from pluto_amd import problems as P


def _spacing(length, shape):
    nz, ny, nx = shape
    return (length[0] / nx, length[1] / ny, length[2] / nz)


@pytest.mark.parametrize("name", ["vecpot_3d_8x7x6.npz", "vecpot_2d_12x10.npz"])
def test_curl_edges_equals_reference_startup(name):
    z = golden(name)
    dims = int(z["dims"][0])
    d = _spacing(z["length"], z["bx"].shape)
    bx, by, bz = P.curl_edges(z["ax"], z["ay"], z["az"], d, dims)
    assert np.abs(z["bx"]).max() > 1.0        # the reference did compute a field
    assert np.array_equal(bx, z["bx"])
    assert np.array_equal(by, z["by"])
    assert np.array_equal(bz, z["bz"])
    cbx, cby, cbz = P.face_average(bx, by, bz, dims)
    assert np.array_equal(cbx, z["prim"][4])
    assert np.array_equal(cby, z["prim"][5])
    assert np.array_equal(cbz, z["prim"][6])


def _wrong_curl(ax, ay, az, d):
    """curl_edges with the y neighbour of A_z taken on the wrong side."""
    dx, dy, dz = d
    bx = (np.roll(az, -1, axis=1) - az) / dy - (ay - np.roll(ay, 1, axis=0)) / dz
    by = -(az - np.roll(az, 1, axis=2)) / dx + (ax - np.roll(ax, 1, axis=0)) / dz
    bz = (ay - np.roll(ay, 1, axis=2)) / dx - (ax - np.roll(ax, 1, axis=1)) / dy
    return bx, by, bz


def test_planted_curl_mistakes_are_caught():
    z = golden("vecpot_3d_8x7x6.npz")
    d = _spacing(z["length"], z["bx"].shape)
    bx, _, _ = _wrong_curl(z["ax"], z["ay"], z["az"], d)
    assert not np.array_equal(bx, z["bx"])
    # a dropped z term of b_x
    assert not np.array_equal((z["az"] - np.roll(z["az"], 1, axis=1)) / d[1], z["bx"])
    # x and z cell widths swapped (the box has dx = dy != dz)
    assert d[0] != d[2]
    bx2, by2, bz2 = P.curl_edges(z["ax"], z["ay"], z["az"], (d[2], d[1], d[0]), 3)
    assert not np.array_equal(bx2, z["bx"])
    assert not np.array_equal(bz2, z["bz"])


def test_curl_edges_live_against_reference():
    from oracle import refrun
    import os
    exe = os.path.join(refrun.REF_DIR, "pluto3d_vecpot")
    if not os.path.exists(exe):
        pytest.skip("reference not built with ASSIGN_VECTOR_POTENTIAL YES (make -C oracle/ref)")
    rng = np.random.default_rng(909)
    shape = (5, 9, 7)
    length = (0.7, 1.3, 0.5)
    prim = np.zeros((8,) + shape)
    prim[0] = 1.0
    prim[7] = 1.0
    zero = np.zeros(shape)
    st = P.MHDState(prim, zero.copy(), zero.copy(), zero.copy(), 3, length)
    apot = tuple(rng.standard_normal(shape) for _ in range(3))
    out = refrun.run(st, 1, dump=[0], fixed_dt=1e-4, vecpot=apot)[0][1]
    bx, by, bz = P.curl_edges(*apot, _spacing(length, shape), 3)
    assert np.array_equal(bx, out.bx) and np.array_equal(by, out.by) and np.array_equal(bz, out.bz)
    assert np.abs(P.divb(out)).max() * min(_spacing(length, shape)) / np.abs(out.bx).max() < 1e-13
