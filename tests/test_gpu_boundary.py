"""Boundary (boundary.c:40-315) on PLUTO's ghosted arrays, against the same
function of the compiled reference (its row library,
oracle/_ref/libplutoref_rows.so) on identical inputs, bit for bit: every
element of Vc and Vs (ghost faces -1 included) after OUTFLOW, REFLECTIVE,
AXISYMMETRIC, EQTSYMMETRIC and PERIODIC sides, mixed per side, internal
sides (type 0) skipped, one direction at a time or all, two and three
ghost zones, non-uniform face areas.  Every array starts random, ghosts
included, so a ghost value read before it is filled (the corner lines of
FillMagneticField) is the same in both.  Unsupported sides change nothing.
"""
import ctypes
import os

import numpy as np
import pytest

from pluto_amd import plutoabi as A
from conftest import ROOT

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "pluto_amd", "libpluto_amd.so")
OUT, REF, AXI, EQT, PER, SHEAR, USER = 1, 2, 3, 4, 5, 6, 7
ALL_DIR = -1


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def libs():
    from oracle import refrun
    from pluto_amd import native as N_
    if N_.device_count() < 1:
        pytest.fail("no HIP device visible")
    path = os.path.join(refrun.REF_DIR, "libplutoref_rows.so")
    if not os.path.exists(path):
        pytest.skip("reference row library not built (make -C oracle/ref)")
    return A.bind_boundary(A.load_abi(LIB)), A.bind_boundary(ctypes.CDLL(path))


class GhostedBox:
    """A PlutoBox whose every element (ghosts and -1 faces included) is
    random, with random positive face areas."""

    def __init__(self, n, ng, bounds, seed):
        from pluto_amd import problems as P
        st = P.random_mhd_3d(max(n), seed=seed)
        st.prim = np.ascontiguousarray(st.prim[:, :n[2], :n[1], :n[0]])
        st.bx, st.by, st.bz = (np.ascontiguousarray(b[:n[2], :n[1], :n[0]])
                               for b in (st.bx, st.by, st.bz))
        self.box = A.PlutoBox(st, ng=ng)
        rng = np.random.default_rng(seed)
        b = self.box
        b.vc[:] = rng.normal(size=b.vc.shape)
        for a in range(3):
            b.vs[a][:] = rng.normal(size=b.vs[a].shape)
        self.area = []
        for a, lo in enumerate(((0, 0, -1), (0, -1, 0), (-1, 0, 0))):
            arr = b.keep(rng.uniform(0.5, 1.5, size=b.vs[a].shape))
            self.area.append(arr)
            b.grid.A[a] = A.table3(b.keep, arr, lo)
        for a in range(3):
            b.grid.lbound[a], b.grid.rbound[a] = bounds[2 * a], bounds[2 * a + 1]

    def arrays(self):
        return [self.box.vc[q].copy() for q in range(8)] + [v.copy() for v in self.box.vs]


def _ref_globals(ref, box):
    g = box.box.grid
    for a, (b, e, n, t) in enumerate(zip(("IBEG", "JBEG", "KBEG"), ("IEND", "JEND", "KEND"),
                                         ("NX1", "NX2", "NX3"), ("NX1_TOT", "NX2_TOT", "NX3_TOT"))):
        ctypes.c_long.in_dll(ref, b).value = g.lbeg[a]
        ctypes.c_long.in_dll(ref, e).value = g.lend[a]
        ctypes.c_long.in_dll(ref, n).value = g.np_int[a]
        ctypes.c_long.in_dll(ref, t).value = g.np_tot[a]


def _run_both(libs, n, ng, bounds, idim, seed):
    mine, ref = libs
    out = []
    for lib in (mine, ref):
        box = GhostedBox(n, ng, bounds, seed)
        if lib is ref:
            _ref_globals(ref, box)
        before = box.arrays()
        lib.Boundary(ctypes.byref(box.box.data), idim, ctypes.byref(box.box.grid))
        out.append((before, box.arrays()))
    return out


NAMES = ["rho", "vx1", "vx2", "vx3", "Bx1", "Bx2", "Bx3", "prs", "bx1s", "bx2s", "bx3s"]

CASES = {
    "outflow": (OUT,) * 6,
    "reflective": (REF,) * 6,
    "axisymmetric": (AXI,) * 6,
    "eqtsymmetric": (EQT,) * 6,
    "periodic": (PER,) * 6,
    "mixed": (OUT, REF, PER, PER, EQT, OUT),
    "mixed_internal": (REF, 0, OUT, AXI, 0, PER),
}


@pytest.mark.parametrize("idim", [ALL_DIR, 0, 1, 2], ids=["all", "x1", "x2", "x3"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_boundary_vs_reference(libs, case, idim):
    n, ng = (9, 8, 7), 2
    _say("case: Boundary %s, idim %d, box %s, %d ghosts" % (case, idim, n, ng))
    (b0, mine), (r0, ref) = _run_both(libs, n, ng, CASES[case], idim, seed=5)
    for q, name in enumerate(NAMES):
        assert np.array_equal(b0[q], r0[q])
        assert np.array_equal(mine[q], ref[q], equal_nan=True), (case, idim, name)
    assert any(not np.array_equal(mine[q], b0[q]) for q in range(11))


@pytest.mark.parametrize("case", ["outflow", "reflective", "mixed"])
def test_boundary_three_ghosts_ragged(libs, case):
    n, ng = (7, 5, 4), 3
    _say("case: Boundary %s, box %s, %d ghosts" % (case, n, ng))
    (_, mine), (_, ref) = _run_both(libs, n, ng, CASES[case], ALL_DIR, seed=9)
    for q, name in enumerate(NAMES):
        assert np.array_equal(mine[q], ref[q], equal_nan=True), (case, name)


def test_outflow_ghosts_divergence_free(libs):
    """Outflow ghosts: the div.B = 0 fill leaves every ghost cell whose six
    faces were set divergence-free (area-weighted), up to round-off."""
    mine, _ = libs
    box = GhostedBox((9, 8, 7), 2, (OUT,) * 6, seed=3)
    mine.Boundary(ctypes.byref(box.box.data), ALL_DIR, ctypes.byref(box.box.grid))
    bx, by, bz = box.box.vs
    ax, ay, az = box.area
    # cell (k, j, i) of the ghosted box: faces i and i-1 at bx[..., i+1], bx[..., i]
    div = (ax[:, :, 1:] * bx[:, :, 1:] - ax[:, :, :-1] * bx[:, :, :-1]) \
        + (ay[:, 1:, :] * by[:, 1:, :] - ay[:, :-1, :] * by[:, :-1, :]) \
        + (az[1:] * bz[1:] - az[:-1] * bz[:-1])
    scale = max(np.abs(v).max() for v in (bx, by, bz))
    # the x1 ghost cells of the last side filled (x3 ghosts fill last, over all i, j)
    assert np.abs(div[:2]).max() <= 1e-12 * scale
    assert np.abs(div[-2:]).max() <= 1e-12 * scale


@pytest.mark.parametrize("bad", [SHEAR, USER])
def test_unsupported_side_changes_nothing(libs, bad):
    # USERDEF needs the program's UserDefBoundary, which this process lacks
    mine, _ = libs
    _say("case: Boundary with an unsupported side type %d" % bad)
    box = GhostedBox((9, 8, 7), 2, (OUT, OUT, bad, OUT, OUT, OUT), seed=2)
    before = box.arrays()
    mine.Boundary(ctypes.byref(box.box.data), ALL_DIR, ctypes.byref(box.box.grid))
    for q in range(11):
        assert np.array_equal(box.arrays()[q], before[q])
    box.box.grid.lbound[1] = box.box.grid.rbound[1] = OUT
    box.box.grid.nproc[0] = 2
    mine.Boundary(ctypes.byref(box.box.data), ALL_DIR, ctypes.byref(box.box.grid))
    for q in range(11):
        assert np.array_equal(box.arrays()[q], before[q])


MAIN_LOOP_BOUNDS = {
    "outflow": ("outflow",) * 6,
    "mixed": ("reflective", "outflow", "periodic", "periodic", "eqtsymmetric", "axisymmetric"),
    # USERDEF sides: oracle/ref/init_io.c's UserDefBoundary (an inflow at
    # X1_BEG that reads the first interior cell, g_intStage and g_time; a
    # zero-gradient copy elsewhere) runs on the host for its side
    "userdef": ("userdef", "outflow", "periodic", "periodic", "reflective", "userdef"),
    "userdef_all": ("userdef",) * 6,
}


@pytest.mark.parametrize("case", sorted(MAIN_LOOP_BOUNDS))
def test_pluto_main_loop_with_library_boundary(libs, case):
    """PLUTO's own main loop and stage with boundary.o replaced by the
    library's Boundary (oracle/_ref/pluto3d_bounddropin): every stage's ghost
    fill runs on the device; states, dt and Mach bit for bit against the
    unmodified reference with the same [Boundary] block."""
    from conftest import assert_states_equal
    from oracle import refrun
    from pluto_amd import problems as P
    exe = os.path.join(refrun.REF_DIR, "pluto3d_bounddropin")
    if not os.path.exists(exe) or not refrun.available(3):
        pytest.skip("Boundary drop-in executable not built (make -C oracle/ref dropin)")
    _say("case: PLUTO main loop with the library's Boundary, %s, 12^3, 6 steps" % case)
    st = P.random_mhd_3d(12, seed=41)
    b = MAIN_LOOP_BOUNDS[case]
    cpu = refrun.run(st, 6, dump=[1, 3], first_dt=1e-4, bounds=b)
    gpu = refrun.run(st, 6, dump=[1, 3], first_dt=1e-4, bounds=b, exe=exe)
    for s in cpu:
        (hc, xc), (hg, xg) = cpu[s], gpu[s]
        assert_states_equal(xg, xc, "%s boundaries, step %d" % (case, s))
        assert hg["dt"] == hc["dt"] and hg["maxmach"] == hc["maxmach"]


ADVANCE_CASES = {
    "outflow": (("outflow",) * 6, "hlld"),
    "mixed": (("reflective", "outflow", "periodic", "periodic", "eqtsymmetric", "axisymmetric"),
              "hlld"),
    "outflow_roe": (("outflow", "reflective", "outflow", "outflow", "reflective", "outflow"),
                    "roe"),
    "userdef": (("userdef", "outflow", "periodic", "periodic", "reflective", "userdef"), "hlld"),
    "userdef_all_hllc": (("userdef",) * 6, "hllc"),
}


@pytest.mark.parametrize("case", sorted(ADVANCE_CASES))
def test_pluto_main_loop_advance_step_non_periodic(libs, case):
    """PLUTO's main loop with AdvanceStep from the library
    (oracle/_ref/pluto3d_dropin) on non-periodic boxes: the ghosted box stays
    on the device for the step, Boundary()'s fill runs there before each
    stage; states, dt and Mach bit for bit against the unmodified
    reference with the same [Boundary] block."""
    from conftest import assert_states_equal
    from oracle import refrun
    from pluto_amd import problems as P
    exe = os.path.join(refrun.REF_DIR, "pluto3d_dropin")
    if not os.path.exists(exe) or not refrun.available(3):
        pytest.skip("drop-in executable not built (make -C oracle/ref dropin)")
    bounds, solver = ADVANCE_CASES[case]
    _say("case: PLUTO main loop with the library's AdvanceStep, %s, %s, 12x10x9, 6 steps"
         % (case, solver))
    big = P.random_mhd_3d(12, seed=47)
    st = big.copy()
    st.prim = np.ascontiguousarray(big.prim[:, :9, :10, :12])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:9, :10, :12]) for b in (big.bx, big.by, big.bz))
    st.length = (12 / 12, 10 / 12, 9 / 12)
    st.prim[4], st.prim[5], st.prim[6] = P.face_average(st.bx, st.by, st.bz)
    cpu = refrun.run(st, 6, dump=[1, 3], first_dt=1e-4, bounds=bounds, solver=solver)
    gpu = refrun.run(st, 6, dump=[1, 3], first_dt=1e-4, bounds=bounds, solver=solver, exe=exe)
    for s in cpu:
        (hc, xc), (hg, xg) = cpu[s], gpu[s]
        assert_states_equal(xg, xc, "%s AdvanceStep, step %d" % (case, s))
        assert hg["dt"] == hc["dt"] and hg["maxmach"] == hc["maxmach"]


STAGE_CASES = {
    "periodic": ("periodic",) * 6,
    "outflow": ("outflow",) * 6,
    "mixed": ("reflective", "outflow", "periodic", "periodic", "eqtsymmetric", "axisymmetric"),
    "userdef": ("userdef", "outflow", "periodic", "periodic", "reflective", "userdef"),
}


@pytest.mark.parametrize("case", sorted(STAGE_CASES))
def test_pluto_main_loop_update_stage(libs, case):
    """PLUTO's own main loop, RK driver and Boundary with update_stage.o
    dropped (oracle/_ref/pluto3d_stagedropin): every stage of every step runs
    through the library's UpdateStage on the ghost zones PLUTO's Boundary
    filled; states, dt and Mach bit for bit against the unmodified
    reference with the same [Boundary] block."""
    from conftest import assert_states_equal
    from oracle import refrun
    from pluto_amd import problems as P
    exe = os.path.join(refrun.REF_DIR, "pluto3d_stagedropin")
    if not os.path.exists(exe) or not refrun.available(3):
        pytest.skip("UpdateStage drop-in executable not built (make -C oracle/ref dropin)")
    bounds = STAGE_CASES[case]
    _say("case: PLUTO main loop with the library's UpdateStage, %s, 12^3, 5 steps" % case)
    st = P.random_mhd_3d(12, seed=61)
    cpu = refrun.run(st, 5, dump=[1, 3], first_dt=1e-4, bounds=bounds)
    gpu = refrun.run(st, 5, dump=[1, 3], first_dt=1e-4, bounds=bounds, exe=exe)
    for s in cpu:
        (hc, xc), (hg, xg) = cpu[s], gpu[s]
        assert_states_equal(xg, xc, "%s UpdateStage, step %d" % (case, s))
        assert hg["dt"] == hc["dt"] and hg["maxmach"] == hc["maxmach"]
