"""The CT entry points (MHD/CT/ct.h:94-112, UCT_HLL) on PLUTO's own arrays,
against the same functions of the compiled reference (its row library,
oracle/_ref/libplutoref_rows.so) on identical inputs, bit for bit:
CT_StoreUpwindEMF / CT_StoreVelSlopes on sweep rows along each direction,
CT_ComputeEMF on a random ghosted box with PLUTO's EMF ranges, CT_Update
with non-uniform cell widths and separate input faces, CT_Allocate, and
PLUTO's own stage linked against the library's CT module.
Arrays start as NaN, so entries outside the ranges must stay untouched.
"""
import ctypes
import os

import numpy as np
import pytest

from pluto_amd import plutoabi as A
from conftest import ROOT

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "pluto_amd", "libpluto_amd.so")
N = (9, 8, 7)          # interior cells (x, y, z)
NG = 2


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
    ref = A.bind_ct(ctypes.CDLL(path))
    ref.ref_rows_max()        # its set-up (NMAX_POINT, state rows)
    return A.load_abi(LIB), ref


TOT = tuple(n + 2 * NG for n in N)     # NX1_TOT, NX2_TOT, NX3_TOT


def _globals(lib, d=0, ijk=(0, 0, 0)):
    for name, v in zip(("NX1_TOT", "NX2_TOT", "NX3_TOT"), TOT):
        ctypes.c_long.in_dll(lib, name).value = v
    ctypes.c_int.in_dll(lib, "g_dir").value = d
    for name, v in zip(("g_i", "g_j", "g_k"), ijk):
        ctypes.c_int.in_dll(lib, name).value = v


class EmfBox:
    """An EMF struct over NaN-filled numpy arrays of the ghosted box."""

    def __init__(self):
        self.keep = A.Keep()
        self.s = A.EMF()
        self.a = {}
        for n in A.EMF_ARRAYS + A.EMF_ARRAYS2:
            arr = self.keep(np.full(TOT[::-1], np.nan))
            self.a[n] = arr
            setattr(self.s, n, A.table3(self.keep, arr))

    def ranges(self):
        return tuple(getattr(self.s, n) for n in ("ibeg", "iend", "jbeg", "jend", "kbeg", "kend"))


def _same(x, y):
    return np.array_equal(x, y, equal_nan=True)


def test_store_upwind_emf_rows(libs):
    mine, ref = libs
    rng = np.random.default_rng(11)
    for d in range(3):
        _say("case: CT_StoreUpwindEMF / CT_StoreVelSlopes, direction %d" % d)
        n = TOT[d]
        rows = A.SweepRows(n + 2)
        rows.flux[:] = rng.normal(size=rows.flux.shape)
        rows.SL[:] = rng.normal(size=rows.SL.shape)
        rows.SR[:] = rng.normal(size=rows.SR.shape)
        rows.vL[:] = rng.normal(size=rows.vL.shape)
        rows.vR[:] = rng.normal(size=rows.vR.shape)
        beg, end = NG - 1, n - NG - 1            # nbeg - 1 .. nend (update_stage.c:196)
        ijk = [3, 4, 2]
        boxes = []
        for lib in (mine, ref):
            _globals(lib, d, ijk)
            e = EmfBox()
            lib.CT_StoreUpwindEMF(ctypes.byref(rows.sw), ctypes.byref(e.s), beg, end, None)
            boxes.append(e)
        a, b = boxes
        for name in A.EMF_ARRAYS + A.EMF_ARRAYS2:
            assert _same(a.a[name], b.a[name]), (d, name)
        assert a.ranges() == b.ranges()


def _random_box(rng, emf, d, vs_keep):
    """Random inputs of CT_ComputeEMF in the arrays of emf and a Data."""
    T1, T2, T3 = TOT
    for n in ("SxL", "SxR", "SyL", "SyR", "SzL", "SzR"):
        emf.a[n][:] = np.abs(rng.normal(size=TOT[::-1]))
    for n in A.EMF_ARRAYS2[15:]:
        emf.a[n][:] = 0.1 * rng.normal(size=TOT[::-1])
    emf.s.ibeg, emf.s.iend = NG - 1, T1 - NG - 1
    emf.s.jbeg, emf.s.jend = NG - 1, T2 - NG - 1
    emf.s.kbeg, emf.s.kend = NG - 1, T3 - NG - 1


def test_compute_emf_and_update(libs):
    from pluto_amd import problems as P
    mine, ref = libs
    rng = np.random.default_rng(12)
    st = P.random_mhd_3d(9, seed=13)
    st.prim = np.ascontiguousarray(st.prim[:, :N[2], :N[1], :N[0]])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:N[2], :N[1], :N[0]]) for b in (st.bx, st.by, st.bz))
    st.length = (N[0] / 9, N[1] / 9, N[2] / 9)
    inputs = None
    out = []
    for lib in (mine, ref):
        _say("case: CT_ComputeEMF + CT_Update, %s" % ("library" if lib is mine else "reference"))
        _globals(lib)
        box = A.PlutoBox(st)
        box.fill_ghosts()
        for a in range(3):                         # fill the face ghosts too
            box.vs[a][np.isnan(box.vs[a])] = 0.0
            box.vs[a][:] += 0.01 * np.arange(box.vs[a].size).reshape(box.vs[a].shape) % 0.37
        e = EmfBox()
        _random_box(np.random.default_rng(12), e, 0, box.keep)
        if inputs is None:
            inputs = [box.vs[a].copy() for a in range(3)]
        for a in range(3):
            box.vs[a][:] = inputs[a]
        box.data.emf = ctypes.addressof(e.s)
        lib.CT_ComputeEMF(ctypes.byref(box.data), ctypes.byref(box.grid))
        # separate input faces and non-uniform widths for the update
        bs = [box.keep(inputs[a] * 1.001) for a in range(3)]
        bs_t = [A.table3(box.keep, bs[0], (0, 0, -1)), A.table3(box.keep, bs[1], (0, -1, 0)),
                A.table3(box.keep, bs[2], (-1, 0, 0))]
        bs_top = box.keep(np.array(bs_t, dtype=np.uint64))
        for a in range(3):
            w = box.keep(np.linspace(0.9, 1.1, TOT[a]) / N[a])
            box.grid.dx[a] = w.ctypes.data
            # face areas: read by the reference's CT_Update, unused in Cartesian
            box.grid.A[a] = A.table3(box.keep, box.keep(np.ones(TOT[::-1])))
        lib.CT_Update(ctypes.byref(box.data), bs_top.ctypes.data, 1.3e-3, ctypes.byref(box.grid))
        out.append((e, box))
    (e0, b0), (e1, b1) = out
    for name in ("ex", "ey", "ez"):
        assert _same(e0.a[name], e1.a[name]), name
        lo = NG - 1
        assert np.isfinite(e0.a[name][lo:TOT[2] - NG, lo:TOT[1] - NG, lo:TOT[0] - NG]).all()
    for a in range(3):
        assert _same(b0.vs[a], b1.vs[a]), "face %d" % a
    assert not _same(b0.vs[0], inputs[0])       # the update did change the faces


def test_allocate(libs):
    mine, _ = libs
    _globals(mine)
    e = A.EMF()
    mine.CT_Allocate(ctypes.byref(e))
    for n in ("ex", "ey", "ez", "ezi", "SxL", "dvz_dz", "dbz_dy"):
        assert getattr(e, n), n
    assert not e.svx


def test_pluto_stage_with_library_ct(libs):
    """PLUTO's own RK2 stage (rk_step.o, update_stage.o) with its CT objects
    replaced by the library's (oracle/_ref/pluto3d_ctdropin) against the
    unmodified reference, bit for bit."""
    from conftest import assert_states_equal
    from oracle import refrun
    from pluto_amd import problems as P
    exe = os.path.join(refrun.REF_DIR, "pluto3d_ctdropin")
    if not os.path.exists(exe) or not refrun.available(3):
        pytest.skip("CT drop-in executable not built (make -C oracle/ref dropin)")
    _say("case: PLUTO main loop with the library's CT module, 12^3, 4 steps")
    st = P.random_mhd_3d(12, seed=43)
    cpu = refrun.run(st, 4, dump=[1, 2], first_dt=1e-4)
    gpu = refrun.run(st, 4, dump=[1, 2], first_dt=1e-4, exe=exe)
    for s in cpu:
        (hc, xc), (hg, xg) = cpu[s], gpu[s]
        assert_states_equal(xg, xc, "CT module step %d" % s)
        assert hg["dt"] == hc["dt"] and hg["maxmach"] == hc["maxmach"]
