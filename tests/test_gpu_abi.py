"""PLUTO's entry points on the GPU, called exactly as a PLUTO build calls them.

* HLLD_Solver, States, PrimToCons, ConsToPrim against the reference's own
  row outputs (tests/golden), bit for bit;
* AdvanceStep on PLUTO-layout arrays (ghost zones, ArrayBox faces) against
  the oracle;
* rk_step.c's sequence rebuilt from UpdateStage, CT_AverageMagneticField,
  PrimToCons and ConsToPrim against the oracle;
* PLUTO's own main loop relinked with this library in place of rk_step.o
  (oracle/_ref/pluto3d_dropin) against the unmodified reference executable.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, assert_states_equal, golden
from pluto_amd import plutoabi as A

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "pluto_amd", "libpluto_amd.so")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def abi():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return A.load_abi(LIB)


def _set_int(lib, name, v):
    ctypes.c_int.in_dll(lib, name).value = v


def _set_dbl(lib, name, v):
    ctypes.c_double.in_dll(lib, name).value = v


def _get_dbl(lib, name):
    return ctypes.c_double.in_dll(lib, name).value


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hlld_solver_vs_reference_rows(abi, d):
    _say("case: HLLD_Solver rows dir %d" % d)
    z = golden("rows_hlld.npz")
    vl, vr = z["hlld%d_vl" % d], z["hlld%d_vr" % d]
    n = vl.shape[0]
    S = A.SweepRows(n)
    S.vL[:n] = vl
    S.vR[1:n + 1] = vr            # stateR.v[i] is storage row i+1
    cmax = np.zeros(n)
    _set_int(abi, "g_dir", d)
    _set_dbl(abi, "g_maxMach", 0.0)
    abi.HLLD_Solver(ctypes.byref(S.sw), 0, n - 1, cmax.ctypes.data, None)
    assert np.array_equal(S.flux[:n], z["hlld%d_flux" % d])
    assert np.array_equal(S.press[:n], z["hlld%d_press" % d])
    assert np.array_equal(S.SL[:n], z["hlld%d_sl" % d])
    assert np.array_equal(S.SR[:n], z["hlld%d_sr" % d])
    assert np.array_equal(cmax, z["hlld%d_cmax" % d])
    assert _get_dbl(abi, "g_maxMach") == z["hlld%d_mach" % d][0]


@pytest.mark.parametrize("d", [0, 1, 2])
def test_states_vs_reference_rows(abi, d):
    _say("case: States rows dir %d" % d)
    z = golden("rows_plm.npz")
    v, bn = z["plm%d_v" % d], z["plm%d_bn" % d]
    n = v.shape[0]
    S = A.SweepRows(n)
    S.vC[:n] = v
    S.bn[:n] = bn
    _set_int(abi, "g_dir", d)
    abi.States(ctypes.byref(S.sw), 1, n - 2, None)
    for i in range(1, n - 1):
        assert np.array_equal(S.vL[i], z["plm%d_vp" % d][i])
        assert np.array_equal(S.stateR_v(i - 1), z["plm%d_vm" % d][i])
        assert np.array_equal(S.uL[i], z["plm%d_up" % d][i])
        assert np.array_equal(S.stateR_u(i - 1), z["plm%d_um" % d][i])


def test_conversions_vs_reference_rows(abi):
    _say("case: PrimToCons / ConsToPrim rows")
    z = golden("rows_conv.npz")
    keep = A.Keep()
    v = keep(z["p2c_v"].copy())
    u = keep(np.zeros_like(v))
    abi.PrimToCons(A.rows_table(keep, v), A.rows_table(keep, u), 0, v.shape[0] - 1)
    assert np.array_equal(u, z["p2c_u"])
    uu = keep(z["c2p_u"].copy())
    vv = keep(np.zeros_like(uu))
    flag = keep(np.zeros(uu.shape[0], dtype=np.uint8))
    rc = abi.ConsToPrim(A.rows_table(keep, uu), A.rows_table(keep, vv), 0, uu.shape[0] - 1,
                        flag.ctypes.data)
    assert rc == 1
    assert np.array_equal(vv, z["c2p_v"])
    assert np.array_equal(uu, z["c2p_uout"])
    assert np.array_equal(flag, z["c2p_flag"])
    # ConsToPrim3D (mappers3D.c:67) passes ibeg = IBEG and the row's flag
    # pointer: flags land at the absolute index of the failing cell
    nr = uu.shape[0]
    for ibeg in (3, 7):
        uu = keep(z["c2p_u"].copy())
        vv = keep(np.zeros_like(uu))
        flag = keep(np.zeros(nr, dtype=np.uint8))
        rc = abi.ConsToPrim(A.rows_table(keep, uu), A.rows_table(keep, vv), ibeg, nr - 1,
                            flag.ctypes.data)
        want = np.zeros(nr, dtype=np.uint8)
        want[ibeg:] = z["c2p_flag"][ibeg:]
        assert rc == int(want.any())
        assert np.array_equal(flag, want)
        assert np.array_equal(vv[ibeg:], z["c2p_v"][ibeg:])
        assert np.array_equal(uu[ibeg:], z["c2p_uout"][ibeg:])
        assert not vv[:ibeg].any()


def _hlld_ptr(abi):
    return ctypes.cast(abi.HLLD_Solver, ctypes.c_void_p).value


def test_advance_step_pluto_layout(abi, oracle_lib):
    from pluto_amd import problems as P
    _say("case: AdvanceStep on PLUTO arrays, 12x10x14 random, 3 steps")
    st = P.random_mhd_3d(14, seed=41)
    st.prim = np.ascontiguousarray(st.prim[:, :, :10, :12])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:, :10, :12]) for b in (st.bx, st.by, st.bz))
    st.length = (12 / 14, 10 / 14, 1.0)
    st.prim[4], st.prim[5], st.prim[6] = P.face_average(st.bx, st.by, st.bz)
    box = A.PlutoBox(st)
    ref = st.copy()
    dt = 1e-3
    assert abi.SetSolver(b"hlld") == _hlld_ptr(abi)
    for s in range(3):
        r = oracle_lib.rk2_step(ref, dt)
        _set_dbl(abi, "g_dt", dt)
        _set_dbl(abi, "g_maxMach", 0.0)
        box.dts.invDt_hyp = 0.0
        rc = abi.AdvanceStep(ctypes.byref(box.data), _hlld_ptr(abi), ctypes.byref(box.dts),
                             ctypes.byref(box.grid))
        assert rc == 0
        assert_states_equal(box.get_state(), ref, "AdvanceStep step %d" % s)
        assert box.dts.invDt_hyp == r["invdt"]
        assert _get_dbl(abi, "g_maxMach") == r["maxmach"]
        dt = oracle_lib.next_dt(r["invdt"], 0.4, 1.1, dt)
        if s == 1:
            abi.pgpu_finalize()      # device memory released; the next call allocates again


def test_advance_step_timing_hook(abi):
    """PGPU_DropinTimes: the periodic AdvanceStep calls since the last reset,
    split into pack, uploads, step, downloads and unpack."""
    from pluto_amd import problems as P

    class Times(ctypes.Structure):
        _fields_ = [(k, ctypes.c_double) for k in
                    ("pack", "h2d", "step", "d2h", "unpack", "total")] + [("calls", ctypes.c_longlong)]

    abi.PGPU_DropinTimes.argtypes = [ctypes.POINTER(Times), ctypes.c_int]
    _say("case: AdvanceStep timing split, 16^3, 3 calls")
    st = P.random_mhd_3d(16, seed=4)
    box = A.PlutoBox(st)
    t = Times()
    abi.PGPU_DropinTimes(ctypes.byref(t), 1)
    _set_dbl(abi, "g_dt", 1e-3)
    for _ in range(3):
        assert abi.AdvanceStep(ctypes.byref(box.data), _hlld_ptr(abi), ctypes.byref(box.dts),
                               ctypes.byref(box.grid)) == 0
    abi.PGPU_DropinTimes(ctypes.byref(t), 1)
    assert t.calls == 3
    for k in ("pack", "h2d", "step", "d2h", "unpack", "total"):
        assert getattr(t, k) > 0.0, k
    assert t.total >= t.step
    abi.PGPU_DropinTimes(ctypes.byref(t), 0)
    assert t.calls == 0 and t.total == 0.0


def test_advance_step_rejects_unsupported(abi):
    from pluto_amd import problems as P
    _say("case: AdvanceStep configuration checks")
    st = P.random_mhd_3d(8, seed=2)
    box = A.PlutoBox(st)
    before = box.vc.copy()
    box.grid.lbound[0] = 2     # not periodic
    rc = abi.AdvanceStep(ctypes.byref(box.data), _hlld_ptr(abi), ctypes.byref(box.dts),
                         ctypes.byref(box.grid))
    assert rc != 0
    assert np.array_equal(box.vc, before, equal_nan=True)
    box.grid.lbound[0] = 5
    rc = abi.AdvanceStep(ctypes.byref(box.data), None, ctypes.byref(box.dts),
                         ctypes.byref(box.grid))
    assert rc != 0


def test_rk_step_from_update_stage(abi, oracle_lib):
    """rk_step.c:85-186 rebuilt from the library's own entry points."""
    from pluto_amd import problems as P
    _say("case: UpdateStage composition, 10^3 random")
    st = P.random_mhd_3d(10, seed=8)
    box = A.PlutoBox(st)
    g = box.ng
    n = 10
    dt = 2e-3
    ref = st.copy()
    r = oracle_lib.rk2_step(ref, dt)
    keep = A.Keep()
    row = keep(np.zeros((n, 8)))
    urow = keep(np.zeros((n, 8)))
    flag = keep(np.zeros(n, dtype=np.uint8))
    rt_v = A.rows_table(keep, row)
    rt_u = A.rows_table(keep, urow)

    def prim_to_cons_3d():
        for k in range(g, g + n):
            for j in range(g, g + n):
                row[:] = box.vc[:, k, j, g:g + n].T
                abi.PrimToCons(rt_v, rt_u, 0, n - 1)
                box.uc[k, j, g:g + n, :] = urow

    def cons_to_prim_3d():
        for k in range(g, g + n):
            for j in range(g, g + n):
                urow[:] = box.uc[k, j, g:g + n, :]
                abi.ConsToPrim(rt_u, rt_v, 0, n - 1, flag.ctypes.data)
                box.uc[k, j, g:g + n, :] = urow
                box.vc[:, k, j, g:g + n] = row.T

    hlld = _hlld_ptr(abi)
    _set_dbl(abi, "g_maxMach", 0.0)
    box.dts.invDt_hyp = 0.0
    # stage 1
    _set_int(abi, "g_intStage", 1)
    prim_to_cons_3d()
    u0 = box.uc.copy()
    b0 = [v.copy() for v in box.vs]
    abi.UpdateStage(ctypes.byref(box.data), box.Uc, None, hlld, dt, ctypes.byref(box.dts),
                    ctypes.byref(box.grid))
    abi.CT_AverageMagneticField(box.Vs, box.Uc, ctypes.byref(box.grid))
    cons_to_prim_3d()
    # stage 2
    _set_int(abi, "g_intStage", 2)
    abi.UpdateStage(ctypes.byref(box.data), box.Uc, None, hlld, dt, ctypes.byref(box.dts),
                    ctypes.byref(box.grid))
    sl = (slice(g, g + n),) * 3
    box.uc[sl] = 0.5 * u0[sl] + 0.5 * box.uc[sl]
    for a in range(3):
        box.vs[a][...] = 0.5 * b0[a] + 0.5 * box.vs[a]
    abi.CT_AverageMagneticField(box.Vs, box.Uc, ctypes.byref(box.grid))
    cons_to_prim_3d()
    assert_states_equal(box.get_state(), ref, "rk_step from UpdateStage")
    assert box.dts.invDt_hyp == r["invdt"]
    assert _get_dbl(abi, "g_maxMach") == r["maxmach"]


def test_pluto_main_loop_with_library(abi):
    from oracle import refrun
    from pluto_amd import problems as P
    exe = os.path.join(refrun.REF_DIR, "pluto3d_dropin")
    if not os.path.exists(exe) or not refrun.available(3):
        pytest.skip("drop-in executable not built (make -C oracle/ref dropin)")
    _say("case: PLUTO main loop linked against libpluto_amd.so, 16^3, 8 CFL steps")
    st = P.random_mhd_3d(16, seed=12)
    cpu = refrun.run(st, 8, dump=[1, 4], first_dt=1e-4)
    gpu = refrun.run(st, 8, dump=[1, 4], first_dt=1e-4, exe=exe)
    for s in cpu:
        (hc, xc), (hg, xg) = cpu[s], gpu[s]
        assert_states_equal(xg, xc, "main loop step %d" % s)
        assert hg["dt"] == hc["dt"] and hg["maxmach"] == hc["maxmach"]
