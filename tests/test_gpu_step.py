"""GPU parity: full RK2 steps of the HIP path against the oracle.

The oracle is the FP64 CPU restatement (itself bit-identical to the compiled
reference, see test_oracle_reference.py / test_oracle_golden.py).  The HIP
library is built without FMA contraction, so the comparisons are bitwise.
"""
import numpy as np
import pytest

from conftest import assert_states_equal, state_diff

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


PIPELINES = pytest.mark.parametrize("pipeline", [0, 1], ids=["fused", "modular"])


@PIPELINES
def test_one_step_3d_fixed_dt(native, oracle_lib, pipeline):
    from pluto_amd import problems as P
    _say("case: 3D random 16^3, 1 step, dt=1e-3")
    st = P.random_mhd_3d(16, seed=3)
    ref = st.copy()
    oracle_lib.rk2_step(ref, 1e-3)
    dom = native.Domain.from_state(st, pipeline=pipeline)
    dom.set_dt(1e-3, fixed=True)
    dom.run(1)
    out = dom.get_state(st)
    assert_states_equal(out, ref, "3D one step")


@PIPELINES
def test_ten_steps_3d_cfl(native, oracle_lib, pipeline):
    from pluto_amd import problems as P
    _say("case: 3D random 16x12x20, 10 CFL steps")
    st = P.random_mhd_3d(16, seed=11)
    # non-cubic box: take a sub-block of a cube, periodicity is not required
    # for parity (both sides wrap the same array)
    st.prim = np.ascontiguousarray(st.prim[:, :, :12, :])
    st.bx = np.ascontiguousarray(st.bx[:, :12, :])
    st.by = np.ascontiguousarray(st.by[:, :12, :])
    st.bz = np.ascontiguousarray(st.bz[:, :12, :])
    from pluto_amd.problems import face_average
    st.prim[4], st.prim[5], st.prim[6] = face_average(st.bx, st.by, st.bz)
    ref = st.copy()
    dts, mach = oracle_lib.run(ref, 10, 1e-4, 0.4, 1.1)
    dom = native.Domain.from_state(st, pipeline=pipeline)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(10)
    out = dom.get_state(st)
    assert_states_equal(out, ref, "3D ten CFL steps")
    gdt, ginv, gmach = dom.history(10)
    assert np.array_equal(gdt, dts)
    assert np.array_equal(gmach, mach)


@PIPELINES
def test_orszag_tang_2d(native, oracle_lib, pipeline):
    from pluto_amd import problems as P
    _say("case: 2D Orszag-Tang 64^2, 40 CFL steps")
    st = P.orszag_tang_2d(64)
    ref = st.copy()
    dts, mach = oracle_lib.run(ref, 40, 1e-4, 0.4, 1.1)
    dom = native.Domain.from_state(st, pipeline=pipeline)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(40)
    out = dom.get_state(st)
    assert_states_equal(out, ref, "2D OT")
    gdt, _, _ = dom.history(40)
    assert np.array_equal(gdt, dts)


@PIPELINES
@pytest.mark.parametrize("shape", [(5, 7, 13), (3, 3, 3), (9, 40, 33), (70, 6, 4)])
def test_ragged_boxes(native, oracle_lib, pipeline, shape):
    """Box sizes that are not multiples of any tile, down to 3 cells a side."""
    from pluto_amd import problems as P
    nz, ny, nx = shape
    _say("case: ragged box %s" % (shape,))
    n = max(shape)
    big = P.random_mhd_3d(n, seed=sum(shape))
    st = big.copy()
    st.prim = np.ascontiguousarray(big.prim[:, :nz, :ny, :nx])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:nz, :ny, :nx]) for b in (big.bx, big.by, big.bz))
    st.length = (nx / n, ny / n, nz / n)
    st.prim[4], st.prim[5], st.prim[6] = P.face_average(st.bx, st.by, st.bz)
    ref = st.copy()
    dts, mach = oracle_lib.run(ref, 3, 1e-4, 0.4, 1.1)
    dom = native.Domain.from_state(st, pipeline=pipeline)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(3)
    assert_states_equal(dom.get_state(st), ref, "box %s" % (shape,))
    gdt, _, gmach = dom.history(3)
    assert np.array_equal(gdt, dts) and np.array_equal(gmach, mach)


@pytest.mark.parametrize("tiling", [(64, 4, 1), (32, 8, 5), (128, 2, 3), (256, 1, 40)])
def test_emf_tiling_does_not_change_results(native, oracle_lib, tiling):
    """The z-marching EMF kernel's workgroup shape and planes per workgroup
    (including segments that end inside the box) leave the step bit-exact."""
    from pluto_amd import problems as P
    _say("case: EMF tiling %s, random 20^3 cut to 11x20x17, 2 CFL steps" % (tiling,))
    big = P.random_mhd_3d(20, seed=77)
    st = big.copy()
    st.prim = np.ascontiguousarray(big.prim[:, :11, :, :17])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:11, :, :17]) for b in (big.bx, big.by, big.bz))
    st.length = (17 / 20, 1.0, 11 / 20)
    st.prim[4], st.prim[5], st.prim[6] = P.face_average(st.bx, st.by, st.bz)
    ref = st.copy()
    oracle_lib.run(ref, 2, 1e-4, 0.4, 1.1)
    native.set_emf_tiling(*tiling)
    try:
        assert native.emf_tiling() == tuple(tiling)
        for pipeline in (0, 1):
            dom = native.Domain.from_state(st, pipeline=pipeline)
            dom.set_dt(1e-4, 0.4, 1.1)
            dom.run(2)
            assert_states_equal(dom.get_state(st), ref, "tiling %s pipeline %d" % (tiling, pipeline))
            dom.close()
    finally:
        native.set_emf_tiling()
    assert native.emf_tiling() == (64, 4, 16)


@pytest.mark.parametrize("kind,kz", [(0, 16), (1, 16), (2, 5), (3, 1), (3, 7), (3, 16)])
@pytest.mark.parametrize("shape", [(18, 20, 34), (5, 3, 4)])
def test_emf_kernels_bit_exact(native, oracle_lib, kind, kz, shape):
    """The register and LDS-staged UCT_HLL EMF kernels, every tile shape and
    planes-per-workgroup count (segments that end inside the box), on boxes
    with nx even that overhang the tiles in x and y and on a 4 x 3 x 5 box
    whose staged halos wrap more than once: bit-exact with the oracle."""
    from pluto_amd import problems as P
    nz, ny, nx = shape
    _say("case: EMF kernel %d kz %d, box %s, 2 CFL steps" % (kind, kz, shape))
    n = max(shape)
    big = P.random_mhd_3d(n, seed=91)
    st = big.copy()
    st.prim = np.ascontiguousarray(big.prim[:, :nz, :ny, :nx])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:nz, :ny, :nx]) for b in (big.bx, big.by, big.bz))
    st.length = (nx / n, ny / n, nz / n)
    st.prim[4], st.prim[5], st.prim[6] = P.face_average(st.bx, st.by, st.bz)
    ref = st.copy()
    oracle_lib.run(ref, 2, 1e-4, 0.4, 1.1)
    native.set_emf_kernel(kind, kz)
    try:
        dom = native.Domain.from_state(st)
        dom.set_dt(1e-4, 0.4, 1.1)
        dom.run(2)
        assert_states_equal(dom.get_state(st), ref, "EMF kernel %d kz %d" % (kind, kz))
        dom.close()
    finally:
        native.set_emf_kernel()


@pytest.mark.parametrize("kind,seg", [(0, 0), (1, 0), (1, 1), (1, 3), (1, 7)])
@pytest.mark.parametrize("shape", [(18, 20, 34), (6, 9, 256), (4, 4, 4), (3, 13, 200)])
def test_flux_xy_kernels_bit_exact(native, oracle_lib, kind, seg, shape):
    """The tile and row-marching x/y flux kernels, y segments that end inside
    the box, rows narrower than a wave (idle lanes), a whole 256-cell row and
    a 4^3 box whose stencils wrap: bit-exact with the oracle."""
    from pluto_amd import problems as P
    nz, ny, nx = shape
    _say("case: flux kernel %d seg %d, box %s, 2 CFL steps" % (kind, seg, shape))
    n = max(shape)
    big = P.random_mhd_3d(n, seed=57)
    st = big.copy()
    st.prim = np.ascontiguousarray(big.prim[:, :nz, :ny, :nx])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:nz, :ny, :nx]) for b in (big.bx, big.by, big.bz))
    st.length = (nx / n, ny / n, nz / n)
    st.prim[4], st.prim[5], st.prim[6] = P.face_average(st.bx, st.by, st.bz)
    ref = st.copy()
    oracle_lib.run(ref, 2, 1e-4, 0.4, 1.1)
    native.set_flux_kernel(kind, seg)
    try:
        dom = native.Domain.from_state(st)
        dom.set_dt(1e-4, 0.4, 1.1)
        dom.run(2)
        assert_states_equal(dom.get_state(st), ref, "flux kernel %d seg %d" % (kind, seg))
        dom.close()
    finally:
        native.set_flux_kernel()


def test_emf_tiling_rejects_bad_shapes(native):
    for bad in [(2, 0), (-1, 0), (1, -1)]:
        with pytest.raises(native.PGPUError):
            native.set_flux_kernel(*bad)
    for bad in [(48, 1, 8), (64, 8, 8), (64, 4, 0)]:
        with pytest.raises(native.PGPUError):
            native.set_emf_tiling(*bad)
    for bad in [(-1, 0), (4, 0), (1, -2)]:
        with pytest.raises(native.PGPUError):
            native.set_emf_kernel(*bad)


@PIPELINES
def test_blast_floors_and_flags(native, oracle_lib, pipeline):
    """Strong blast: every HLLD region; flags and floors as the oracle."""
    from pluto_amd import problems as P
    _say("case: blast 24^3, 6 CFL steps")
    st = P.blast_3d(24, p_in=1e4, p_out=1e-4)
    ref = st.copy()
    dom = native.Domain.from_state(st, pipeline=pipeline)
    dom.set_dt(1e-6, 0.4, 1.1)
    dt = 1e-6
    for s in range(6):
        r = oracle_lib.rk2_step(ref, dt)
        dt = oracle_lib.next_dt(r["invdt"], 0.4, 1.1, dt)
        dom.run(1)
        assert np.array_equal(dom.flags(), r["flag"])
    assert_states_equal(dom.get_state(st), ref, "blast")


def _divb_rel(st):
    from pluto_amd import problems as P
    db = P.divb(st)
    bmax = max(np.abs(st.bx).max(), np.abs(st.by).max(), np.abs(st.bz).max())
    return float(np.abs(db).max() * min(st.spacing()) / bmax)


def test_orszag_tang_128_200_steps(native, oracle_lib):
    """The 2D Orszag-Tang workload at full size: 128^2, 200 CFL steps (CFL 0.4,
    first dt 1e-4, CFL_max_var 1.1), identical to the oracle (so far inside
    the 1e-12 per step / 1e-9 after 100 steps bar), div B at round-off."""
    from pluto_amd import problems as P
    _say("case: Orszag-Tang 128^2, 200 CFL steps")
    st = P.orszag_tang_2d(128)
    ref = st.copy()
    dts, mach = oracle_lib.run(ref, 200, 1e-4, 0.4, 1.1)
    dom = native.Domain.from_state(st)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(200)
    out = dom.get_state(st)
    assert_states_equal(out, ref, "OT 128^2 after 200 steps")
    gdt, _, _ = dom.history(200)
    assert np.array_equal(gdt, dts)
    assert _divb_rel(out) <= 1e-13


def test_random_3d_100_steps(native, oracle_lib):
    """100 CFL steps of the 3D random-field state (48^3): identical to the
    oracle after every step, div B at round-off."""
    from pluto_amd import problems as P
    _say("case: random 3D 48^3, 100 CFL steps")
    st = P.random_mhd_3d(48, seed=2024)
    ref = st.copy()
    dts, mach = oracle_lib.run(ref, 100, 1e-4, 0.4, 1.1)
    dom = native.Domain.from_state(st)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(100)
    out = dom.get_state(st)
    assert_states_equal(out, ref, "48^3 after 100 steps")
    gdt, _, gmach = dom.history(100)
    assert np.array_equal(gdt, dts) and np.array_equal(gmach, mach)
    assert _divb_rel(out) <= 1e-13
