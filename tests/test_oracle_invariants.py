"""Properties the mathematics fixes, pinning the oracle beyond the vectors.

Each check is chosen so a plausible restatement error (dropped term, wrong
sign or index, transposed operand) fails it:
  * translation invariance on the periodic box is exact (bitwise);
  * a uniform state stays uniform;
  * mass, momentum and energy are conserved to round-off;
  * the discrete divergence of the staggered field stays at round-off
    (ct_update.c:281-313 defines it);
  * HLLD returns the physical flux (textbook ideal-MHD flux) for identical
    states and for supersonic flow, where every HLL-type solver must;
  * mirror and axis-permutation symmetry hold to round-off.
"""
import math

import numpy as np
import pytest

from conftest import max_rel_diff
from pluto_amd import problems as P

GAMMA = 5.0 / 3.0


def totals(st):
    rho = st.prim[0]
    m = rho * st.prim[1:4]
    b2 = st.prim[4] ** 2 + st.prim[5] ** 2 + st.prim[6] ** 2
    e = 0.5 * rho * np.sum(st.prim[1:4] ** 2, axis=0) + 0.5 * b2 + st.prim[7] / (GAMMA - 1.0)
    return np.array([rho.sum(), m[0].sum(), m[1].sum(), m[2].sum(), e.sum()])


def test_translation_exact(oracle_lib):
    st = P.random_mhd_3d(12, seed=5)
    sh = (3, 5, 7)
    rolled = P.MHDState(np.roll(st.prim, sh, axis=(1, 2, 3)), np.roll(st.bx, sh, axis=(0, 1, 2)),
                        np.roll(st.by, sh, axis=(0, 1, 2)), np.roll(st.bz, sh, axis=(0, 1, 2)))
    a, b = st.copy(), rolled.copy()
    oracle_lib.run(a, 3, 2e-3, fixed_dt=True)
    oracle_lib.run(b, 3, 2e-3, fixed_dt=True)
    assert np.array_equal(np.roll(a.prim, sh, axis=(1, 2, 3)), b.prim)
    assert np.array_equal(np.roll(a.bx, sh, axis=(0, 1, 2)), b.bx)


def test_uniform_state_stays_uniform(oracle_lib):
    st = P.uniform_state((6, 5, 7), rho=1.3, v=(0.2, -0.1, 0.3), b=(0.5, 0.2, -0.4), p=0.7)
    st0 = st.copy()
    oracle_lib.run(st, 4, 1e-2, fixed_dt=True)
    for q in range(8):
        assert np.ptp(st.prim[q]) == 0.0
    assert max_rel_diff(st, st0) < 1e-14


def test_conservation_and_divb(oracle_lib):
    st = P.random_mhd_3d(16, seed=9)
    t0 = totals(st)
    oracle_lib.run(st, 10, 1e-3, 0.4, 1.1)
    t1 = totals(st)
    scale = np.array([abs(t0[0]), np.abs(t0[1:4]).max() + t0[0], 0, 0, abs(t0[4])])
    scale[2] = scale[3] = scale[1]
    assert np.all(np.abs(t1 - t0) <= 1e-12 * scale), (t1 - t0) / scale
    dx = 1.0 / 16
    bmax = max(np.abs(st.bx).max(), np.abs(st.by).max(), np.abs(st.bz).max())
    assert np.abs(P.divb(st)).max() * dx / bmax < 1e-13


def mhd_flux(v, d):
    """Textbook ideal-MHD flux along axis d of a primitive state."""
    rho, vel, b, p = v[0], v[1:4], v[4:7], v[7]
    b2 = b @ b
    pt = p + 0.5 * b2
    e = 0.5 * rho * (vel @ vel) + 0.5 * b2 + p / (GAMMA - 1.0)
    f = np.zeros(8)
    f[0] = rho * vel[d]
    f[1:4] = rho * vel[d] * vel - b[d] * b
    f[1 + d] += pt                      # total pressure on the normal momentum
    f[4:7] = vel[d] * b - b[d] * vel
    f[7] = (e + pt) * vel[d] - b[d] * (vel @ b)
    return f


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hlld_physical_flux_limits(oracle_lib, d):
    rng = np.random.default_rng(3 + d)
    n = 64
    v = np.empty((n, 8))
    v[:, 0] = rng.uniform(0.5, 2, n)
    v[:, 1:4] = rng.normal(0, 0.5, (n, 3))
    v[:, 4:7] = rng.normal(0, 1, (n, 3))
    v[:, 7] = rng.uniform(0.5, 2, n)
    # identical states
    flux, press, sl, sr, cmax, _ = oracle_lib.hlld_rows(d, v, v)
    for i in range(n):
        f = flux[i].copy()
        f[1 + d] += press[i]          # the solver keeps total pressure apart
        ref = mhd_flux(v[i], d)
        ref[4 + d] = 0.0
        assert np.allclose(f, ref, rtol=1e-12, atol=1e-12)
    # supersonic to the right: the left state's flux
    vl, vr = v.copy(), v[::-1].copy()
    vl[:, 1 + d] += 30.0
    vr[:, 1 + d] += 30.0
    vr[:, 4 + d] = vl[:, 4 + d]
    flux, press, sl, sr, cmax, _ = oracle_lib.hlld_rows(d, vl, vr)
    assert np.all(sl >= 0)
    for i in range(n):
        f = flux[i].copy()
        f[1 + d] += press[i]
        ref = mhd_flux(vl[i], d)
        ref[4 + d] = 0.0
        assert np.allclose(f, ref, rtol=1e-12, atol=1e-12)


def test_mirror_symmetry(oracle_lib):
    st = P.random_mhd_3d(12, seed=21)
    m = st.copy()
    # x -> -x: cells reverse, vx and Bx change sign; x-faces move to the other side
    m.prim = np.ascontiguousarray(st.prim[:, :, :, ::-1])
    m.prim[1] *= -1
    m.prim[4] *= -1
    m.bx = -np.ascontiguousarray(np.roll(st.bx[:, :, ::-1], -1, axis=2))
    m.by = np.ascontiguousarray(st.by[:, :, ::-1])
    m.bz = np.ascontiguousarray(st.bz[:, :, ::-1])
    cbx, cby, cbz = P.face_average(m.bx, m.by, m.bz)
    assert np.allclose(cbx, m.prim[4], rtol=0, atol=1e-14)
    a, b = st.copy(), m.copy()
    oracle_lib.run(a, 3, 2e-3, fixed_dt=True)
    oracle_lib.run(b, 3, 2e-3, fixed_dt=True)
    back = np.ascontiguousarray(b.prim[:, :, :, ::-1])
    back[1] *= -1
    back[4] *= -1
    scale = np.abs(a.prim).max(axis=(1, 2, 3))
    err = np.abs(back - a.prim).max(axis=(1, 2, 3)) / scale
    assert err.max() < 1e-12, err


def test_axis_permutation(oracle_lib):
    st = P.random_mhd_3d(10, seed=33)
    # relabel (x, y, z) -> (y, z, x): array axes (z, y, x) -> (x, z, y)
    t = st.copy()
    t.prim = np.ascontiguousarray(np.transpose(st.prim, (0, 3, 1, 2)))
    t.prim = t.prim[[0, 2, 3, 1, 5, 6, 4, 7]]
    t.bx = np.ascontiguousarray(np.transpose(st.by, (2, 0, 1)))
    t.by = np.ascontiguousarray(np.transpose(st.bz, (2, 0, 1)))
    t.bz = np.ascontiguousarray(np.transpose(st.bx, (2, 0, 1)))
    a, b = st.copy(), t.copy()
    oracle_lib.run(a, 2, 2e-3, fixed_dt=True)
    oracle_lib.run(b, 2, 2e-3, fixed_dt=True)
    back = np.transpose(b.prim[[0, 3, 1, 2, 6, 4, 5, 7]], (0, 2, 3, 1))
    scale = np.abs(a.prim).max(axis=(1, 2, 3))
    err = np.abs(back - a.prim).max(axis=(1, 2, 3)) / scale
    assert err.max() < 1e-12, err
