"""The HLL, HLLC, Lax-Friedrichs and Roe Riemann solvers behind the same
pointers (PLUTO's ``Solver hll`` / ``hllc`` / ``tvdlf`` / ``roe``).

Row level: the library's HLL_Solver / HLLC_Solver / LF_Solver / Roe_Solver
entry points against the reference's own rows (tests/golden/rows_<solver>.npz).
Box level: domains with PGPU_SOLVER_HLL / _HLLC / _TVDLF / _ROE
against PLUTO's main loop run with that Solver (tests/golden/step_*_<solver>.npz)
and against the oracle on ragged boxes (odd rows take the x/y tile kernel,
even ones the row kernel),
both pipelines, bit for bit; AdvanceStep called with the pointer SetSolver
returns; PLUTO's main loop linked against the library.  FP32 flux stays
within 1e-6 (L1) of FP64.
"""
import ctypes
import os

import numpy as np
import pytest

from pluto_amd import plutoabi as A
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


SOLVERS = pytest.mark.parametrize("solver", ["hll", "hllc", "tvdlf", "roe"])
ENTRY = {"hll": "HLL_Solver", "hllc": "HLLC_Solver", "tvdlf": "LF_Solver", "roe": "Roe_Solver"}


def _code(native, solver):
    return {"hll": native.SOLVER_HLL, "hllc": native.SOLVER_HLLC, "tvdlf": native.SOLVER_TVDLF,
            "roe": native.SOLVER_ROE}[solver]


@SOLVERS
@pytest.mark.parametrize("d", [0, 1, 2])
def test_solver_rows_vs_reference(abi, solver, d):
    _say("case: %s rows dir %d" % (solver, d))
    z = golden("rows_%s.npz" % solver)
    k = solver
    vl, vr = z["%s%d_vl" % (k, d)], z["%s%d_vr" % (k, d)]
    n = vl.shape[0]
    S = A.SweepRows(n)
    S.vL[:n] = vl
    S.vR[1:n + 1] = vr
    cmax = np.zeros(n)
    ctypes.c_int.in_dll(abi, "g_dir").value = d
    ctypes.c_double.in_dll(abi, "g_maxMach").value = 0.0
    fn = getattr(abi, ENTRY[solver])
    fn(ctypes.byref(S.sw), 0, n - 1, cmax.ctypes.data, None)
    assert np.array_equal(S.flux[:n], z["%s%d_flux" % (k, d)])
    assert np.array_equal(S.press[:n], z["%s%d_press" % (k, d)])
    assert np.array_equal(S.SL[:n], z["%s%d_sl" % (k, d)])
    assert np.array_equal(S.SR[:n], z["%s%d_sr" % (k, d)])
    assert np.array_equal(cmax, z["%s%d_cmax" % (k, d)])
    assert ctypes.c_double.in_dll(abi, "g_maxMach").value == z["%s%d_mach" % (k, d)][0]


@SOLVERS
@pytest.mark.parametrize("pipeline", [0, 1], ids=["fused", "modular"])
@pytest.mark.parametrize("case", ["step_3d_8", "step_ot2d_32"])
def test_domain_vs_reference_run(native, solver, case, pipeline):
    name = "%s_%s.npz" % (case, solver)
    _say("case: %s domain vs PLUTO Solver %s, %s, pipeline %d" % (solver, solver, name, pipeline))
    init, steps, run = load_run(name)
    dom = native.Domain.from_state(init, pipeline=pipeline, solver=_code(native, solver))
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


@SOLVERS
@pytest.mark.parametrize("shape", [(5, 7, 13), (9, 40, 33), (6, 11, 22)])
def test_ragged_boxes_vs_oracle(native, oracle_lib, solver, shape):
    from pluto_amd import problems as P
    nz, ny, nx = shape
    _say("case: %s ragged box %s" % (solver, shape))
    n = max(shape)
    big = P.random_mhd_3d(n, seed=7 + sum(shape))
    st = big.copy()
    st.prim = np.ascontiguousarray(big.prim[:, :nz, :ny, :nx])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:nz, :ny, :nx]) for b in (big.bx, big.by, big.bz))
    st.length = (nx / n, ny / n, nz / n)
    st.prim[4], st.prim[5], st.prim[6] = P.face_average(st.bx, st.by, st.bz)
    ref = st.copy()
    dts, mach = oracle_lib.run(ref, 3, 1e-4, 0.4, 1.1, solver=solver)
    dom = native.Domain.from_state(st, solver=_code(native, solver))
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(3)
    assert_states_equal(dom.get_state(st), ref, "%s box %s" % (solver, shape))
    gdt, _, gmach = dom.history(3)
    assert np.array_equal(gdt, dts) and np.array_equal(gmach, mach)


@SOLVERS
def test_advance_step_with_solver_pointer(abi, oracle_lib, solver):
    from pluto_amd import problems as P
    _say("case: AdvanceStep with SetSolver(%s), 12^3 random, 2 steps" % solver)
    st = P.random_mhd_3d(12, seed=19)
    box = A.PlutoBox(st)
    ref = st.copy()
    ptr = abi.SetSolver(solver.encode())
    fn = getattr(abi, ENTRY[solver])
    assert ptr == ctypes.cast(fn, ctypes.c_void_p).value
    dt = 1e-3
    for s in range(2):
        r = oracle_lib.rk2_step(ref, dt, solver=solver)
        ctypes.c_double.in_dll(abi, "g_dt").value = dt
        ctypes.c_double.in_dll(abi, "g_maxMach").value = 0.0
        box.dts.invDt_hyp = 0.0
        rc = abi.AdvanceStep(ctypes.byref(box.data), ptr, ctypes.byref(box.dts),
                             ctypes.byref(box.grid))
        assert rc == 0
        assert_states_equal(box.get_state(), ref, "AdvanceStep (%s) step %d" % (solver, s))
        assert box.dts.invDt_hyp == r["invdt"]
        assert ctypes.c_double.in_dll(abi, "g_maxMach").value == r["maxmach"]
        dt = oracle_lib.next_dt(r["invdt"], 0.4, 1.1, dt)


@SOLVERS
def test_pluto_main_loop_with_library(native, solver):
    from oracle import refrun
    from pluto_amd import problems as P
    exe = os.path.join(refrun.REF_DIR, "pluto3d_dropin")
    if not os.path.exists(exe) or not refrun.available(3):
        pytest.skip("drop-in executable not built (make -C oracle/ref dropin)")
    _say("case: PLUTO main loop with Solver %s linked against libpluto_amd.so, 12^3, 6 steps" % solver)
    st = P.random_mhd_3d(12, seed=23)
    cpu = refrun.run(st, 6, dump=[1, 3], first_dt=1e-4, solver=solver)
    gpu = refrun.run(st, 6, dump=[1, 3], first_dt=1e-4, exe=exe, solver=solver)
    for s in cpu:
        (hc, xc), (hg, xg) = cpu[s], gpu[s]
        assert_states_equal(xg, xc, "main loop (%s) step %d" % (solver, s))
        assert hg["dt"] == hc["dt"] and hg["maxmach"] == hc["maxmach"]


@SOLVERS
def test_fp32_flux_close_to_fp64(native, solver):
    from pluto_amd import problems as P
    _say("case: %s FP32 flux vs FP64, 24^3, 10 steps" % solver)
    st = P.random_mhd_3d(24, seed=5)
    a = native.Domain.from_state(st, solver=_code(native, solver))
    b = native.Domain.from_state(st, solver=_code(native, solver), precision=native.PGPU_PREC_FP32_FLUX)
    for d in (a, b):
        d.set_dt(1e-4, 0.4, 1.1)
        d.run(10)
    err = b.error_vs(a)
    assert max(v[0] for v in err.values()) < 1e-6, err
    a.close()
    b.close()


@SOLVERS
def test_adjoint_rejects_other_solvers(native, solver):
    import torch
    from pluto_amd import problems as P
    st = P.random_mhd_3d(8, seed=1)
    dom = native.Domain.from_state(st, solver=_code(native, solver))
    dom.set_dt(1e-4, fixed=True)
    lp = torch.zeros((8,) + st.shape, dtype=torch.float64, device="cuda")
    lf = torch.zeros((3,) + st.shape, dtype=torch.float64, device="cuda")
    p8 = ctypes.c_void_p(lp.data_ptr())
    f3 = ctypes.c_void_p(lf.data_ptr())
    rc = native.lib().pgpu_adjoint_step(dom.h, p8, f3)
    assert rc != 0
    dom.close()
