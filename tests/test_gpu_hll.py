"""The HLL Riemann solver behind the same pointers (PLUTO's ``Solver hll``).

Row level: PLUTO's HLL_Solver entry point of the library against the
reference's own HLL_Solver rows (tests/golden/rows_hll.npz).  Box level: a
domain with PGPU_SOLVER_HLL against PLUTO's main loop run with Solver hll
(tests/golden/step_*_hll.npz) and against the oracle on ragged boxes, both
pipelines, bit for bit; AdvanceStep called with the HLL_Solver pointer that
SetSolver("hll") returns; PLUTO's main loop linked against the library with
Solver hll.  FP32 flux with HLL stays within 1e-6 (L1) of FP64.
"""
import ctypes
import os

import numpy as np
import pytest

import plutoabi as A
from conftest import ROOT, assert_states_equal, golden, load_run

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "pluto_amd", "libpluto_amd.so")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


@pytest.fixture(scope="module")
def abi(native):
    return A.load_abi(LIB)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hll_solver_rows_vs_reference(abi, d):
    _say("case: HLL_Solver rows dir %d" % d)
    z = golden("rows_hll.npz")
    vl, vr = z["hll%d_vl" % d], z["hll%d_vr" % d]
    n = vl.shape[0]
    S = A.SweepRows(n)
    S.vL[:n] = vl
    S.vR[1:n + 1] = vr
    cmax = np.zeros(n)
    ctypes.c_int.in_dll(abi, "g_dir").value = d
    ctypes.c_double.in_dll(abi, "g_maxMach").value = 0.0
    abi.HLL_Solver(ctypes.byref(S.sw), 0, n - 1, cmax.ctypes.data, None)
    assert np.array_equal(S.flux[:n], z["hll%d_flux" % d])
    assert np.array_equal(S.press[:n], z["hll%d_press" % d])
    assert np.array_equal(S.SL[:n], z["hll%d_sl" % d])
    assert np.array_equal(S.SR[:n], z["hll%d_sr" % d])
    assert np.array_equal(cmax, z["hll%d_cmax" % d])
    assert ctypes.c_double.in_dll(abi, "g_maxMach").value == z["hll%d_mach" % d][0]


@pytest.mark.parametrize("pipeline", [0, 1], ids=["fused", "modular"])
@pytest.mark.parametrize("name", ["step_3d_8_hll.npz", "step_ot2d_32_hll.npz"])
def test_domain_hll_vs_reference_run(native, name, pipeline):
    _say("case: HLL domain vs PLUTO Solver hll, %s, pipeline %d" % (name, pipeline))
    init, steps, run = load_run(name)
    dom = native.Domain.from_state(init, pipeline=pipeline, solver=native.SOLVER_HLL)
    dom.set_dt(run["first_dt"], run["cfl"], run["cfl_max_var"])
    done = 0
    for s in sorted(steps):
        dom.run(s - done)
        done = s
        h, ref = steps[s]
        assert_states_equal(dom.get_state(init), ref, "%s step %d" % (name, s))
        sc = dom.scalars()
        assert sc["dt"] == h["dt"] and sc["maxmach"] == h["maxmach"]
    dom.close()


@pytest.mark.parametrize("shape", [(5, 7, 13), (9, 40, 33)])
def test_hll_ragged_boxes_vs_oracle(native, oracle_lib, shape):
    from pluto_amd import problems as P
    nz, ny, nx = shape
    _say("case: HLL ragged box %s" % (shape,))
    n = max(shape)
    big = P.random_mhd_3d(n, seed=7 + sum(shape))
    st = big.copy()
    st.prim = np.ascontiguousarray(big.prim[:, :nz, :ny, :nx])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:nz, :ny, :nx]) for b in (big.bx, big.by, big.bz))
    st.length = (nx / n, ny / n, nz / n)
    st.prim[4], st.prim[5], st.prim[6] = P.face_average(st.bx, st.by, st.bz)
    ref = st.copy()
    dts, mach = oracle_lib.run(ref, 3, 1e-4, 0.4, 1.1, solver="hll")
    dom = native.Domain.from_state(st, solver=native.SOLVER_HLL)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(3)
    assert_states_equal(dom.get_state(st), ref, "HLL box %s" % (shape,))
    gdt, _, gmach = dom.history(3)
    assert np.array_equal(gdt, dts) and np.array_equal(gmach, mach)


def test_advance_step_with_hll_pointer(abi, oracle_lib):
    from pluto_amd import problems as P
    _say("case: AdvanceStep with SetSolver(\"hll\"), 12^3 random, 2 steps")
    st = P.random_mhd_3d(12, seed=19)
    box = A.PlutoBox(st)
    ref = st.copy()
    ptr = abi.SetSolver(b"hll")
    assert ptr == ctypes.cast(abi.HLL_Solver, ctypes.c_void_p).value
    dt = 1e-3
    for s in range(2):
        r = oracle_lib.rk2_step(ref, dt, solver="hll")
        ctypes.c_double.in_dll(abi, "g_dt").value = dt
        ctypes.c_double.in_dll(abi, "g_maxMach").value = 0.0
        box.dts.invDt_hyp = 0.0
        rc = abi.AdvanceStep(ctypes.byref(box.data), ptr, ctypes.byref(box.dts),
                             ctypes.byref(box.grid))
        assert rc == 0
        assert_states_equal(box.get_state(), ref, "AdvanceStep (HLL) step %d" % s)
        assert box.dts.invDt_hyp == r["invdt"]
        assert ctypes.c_double.in_dll(abi, "g_maxMach").value == r["maxmach"]
        dt = oracle_lib.next_dt(r["invdt"], 0.4, 1.1, dt)


def test_pluto_main_loop_with_library_hll(native):
    from oracle import refrun
    from pluto_amd import problems as P
    exe = os.path.join(refrun.REF_DIR, "pluto3d_dropin")
    if not os.path.exists(exe) or not refrun.available(3):
        pytest.skip("drop-in executable not built (make -C oracle/ref dropin)")
    _say("case: PLUTO main loop with Solver hll linked against libpluto_amd.so, 12^3, 6 steps")
    st = P.random_mhd_3d(12, seed=23)
    cpu = refrun.run(st, 6, dump=[1, 3], first_dt=1e-4, solver="hll")
    gpu = refrun.run(st, 6, dump=[1, 3], first_dt=1e-4, exe=exe, solver="hll")
    for s in cpu:
        (hc, xc), (hg, xg) = cpu[s], gpu[s]
        assert_states_equal(xg, xc, "main loop (HLL) step %d" % s)
        assert hg["dt"] == hc["dt"] and hg["maxmach"] == hc["maxmach"]


def test_hll_fp32_flux_close_to_fp64(native):
    from pluto_amd import problems as P
    _say("case: HLL FP32 flux vs FP64, 24^3, 10 steps")
    st = P.random_mhd_3d(24, seed=5)
    a = native.Domain.from_state(st, solver=native.SOLVER_HLL)
    b = native.Domain.from_state(st, solver=native.SOLVER_HLL, precision=native.PGPU_PREC_FP32_FLUX)
    for d in (a, b):
        d.set_dt(1e-4, 0.4, 1.1)
        d.run(10)
    err = b.error_vs(a)
    assert max(v[0] for v in err.values()) < 1e-6, err
    a.close()
    b.close()


def test_adjoint_rejects_hll(native):
    import torch
    from pluto_amd import problems as P
    st = P.random_mhd_3d(8, seed=1)
    dom = native.Domain.from_state(st, solver=native.SOLVER_HLL)
    dom.set_dt(1e-4, fixed=True)
    lp = torch.zeros((8,) + st.shape, dtype=torch.float64, device="cuda")
    lf = torch.zeros((3,) + st.shape, dtype=torch.float64, device="cuda")
    p8 = ctypes.c_void_p(lp.data_ptr())
    f3 = ctypes.c_void_p(lf.data_ptr())
    rc = native.lib().pgpu_adjoint_step(dom.h, p8, f3)
    assert rc != 0
    dom.close()
