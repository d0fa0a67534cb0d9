"""The ARITHMETIC, UCT0 and UCT_CONTACT EMF averages (PLUTO's CT_EMF_AVERAGE
other than the default UCT_HLL) on the fused FP64 pipeline.

Box level: domains with each average against PLUTO's main loop built with
that CT_EMF_AVERAGE (tests/golden/step_*_<average>.npz: a random 3D field, a
box at rest with a moving blob, whose zero mass fluxes take UCT_CONTACT's
third branch, and Orszag-Tang 2D), bit for bit, states, dt and Mach; the
other Riemann solvers with each average against the oracle on ragged boxes;
AdvanceStep after PGPU_SetEMFAverage; PLUTO's main loop built with each
average and linked against the library.  Configurations that keep the
UCT_HLL data only (FP32 flux, modular pipeline, adjoint, UpdateStage) refuse
the other averages.
"""
import ctypes
import os

import numpy as np
import pytest

from pluto_amd import plutoabi as A
from conftest import ROOT, assert_states_equal, load_run

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "pluto_amd", "libpluto_amd.so")
AVERAGES = ["arithmetic", "uct0", "uct_contact"]


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


@pytest.mark.parametrize("emf", AVERAGES)
@pytest.mark.parametrize("case", ["step_3d_8", "step_blob_12x10x8", "step_ot2d_32"])
def test_domain_vs_reference_run(native, emf, case):
    name = "%s_%s.npz" % (case, emf)
    _say("case: %s domain vs PLUTO's main loop, %s" % (emf, name))
    init, steps, run = load_run(name)
    dom = native.Domain.from_state(init, emf_average=native.EMF_AVERAGES[emf])
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


@pytest.mark.parametrize("emf", AVERAGES)
@pytest.mark.parametrize("solver", ["hlld", "hll", "hllc", "tvdlf", "roe"])
def test_solvers_on_ragged_box_vs_oracle(native, oracle_lib, emf, solver):
    from pluto_amd import problems as P
    nz, ny, nx = shape = (9, 21, 35)
    _say("case: %s with %s, ragged box %s, 3 CFL steps" % (emf, solver, shape))
    big = P.random_mhd_3d(36, seed=11)
    st = big.copy()
    st.prim = np.ascontiguousarray(big.prim[:, :nz, :ny, :nx])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:nz, :ny, :nx]) for b in (big.bx, big.by, big.bz))
    st.length = (nx / 36, ny / 36, nz / 36)
    st.prim[4], st.prim[5], st.prim[6] = P.face_average(st.bx, st.by, st.bz)
    ref = st.copy()
    dts, mach = oracle_lib.run(ref, 3, 1e-4, 0.4, 1.1, solver=solver, emf=emf)
    code = {"hlld": native.SOLVER_HLLD, "hll": native.SOLVER_HLL, "hllc": native.SOLVER_HLLC,
            "tvdlf": native.SOLVER_TVDLF, "roe": native.SOLVER_ROE}
    dom = native.Domain.from_state(st, solver=code[solver], emf_average=native.EMF_AVERAGES[emf])
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(3)
    assert_states_equal(dom.get_state(st), ref, "%s/%s box %s" % (emf, solver, shape))
    gdt, _, gmach = dom.history(3)
    assert np.array_equal(gdt, dts) and np.array_equal(gmach, mach)
    dom.close()


@pytest.mark.parametrize("emf", AVERAGES)
def test_large_box_divb(native, oracle_lib, emf):
    """64^3, 5 steps: bit-exact against the oracle and div.B at round-off."""
    from pluto_amd import problems as P
    _say("case: %s 64^3 random field, 5 CFL steps" % emf)
    st = P.random_mhd_3d(64, seed=31)
    ref = st.copy()
    oracle_lib.run(ref, 5, 1e-4, 0.4, 1.1, emf=emf)
    dom = native.Domain.from_state(st, emf_average=native.EMF_AVERAGES[emf])
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(5)
    out = dom.get_state(st)
    assert_states_equal(out, ref, "%s 64^3" % emf)
    bmax = max(np.abs(out.bx).max(), np.abs(out.by).max(), np.abs(out.bz).max())
    assert np.abs(P.divb(out)).max() * out.spacing()[0] / bmax <= 1e-13
    dom.close()


@pytest.mark.parametrize("emf", AVERAGES)
def test_advance_step_after_set_emf_average(abi, oracle_lib, emf):
    from oracle.oracle import EMF_AVERAGES
    from pluto_amd import problems as P
    _say("case: AdvanceStep after PGPU_SetEMFAverage(%s), 12^3 random, 2 steps" % emf)
    st = P.random_mhd_3d(12, seed=29)
    box = A.PlutoBox(st)
    ref = st.copy()
    ptr = abi.SetSolver(b"hlld")
    assert abi.PGPU_SetEMFAverage(EMF_AVERAGES[emf]) == 0
    try:
        dt = 1e-3
        for s in range(2):
            r = oracle_lib.rk2_step(ref, dt, emf=emf)
            ctypes.c_double.in_dll(abi, "g_dt").value = dt
            ctypes.c_double.in_dll(abi, "g_maxMach").value = 0.0
            box.dts.invDt_hyp = 0.0
            rc = abi.AdvanceStep(ctypes.byref(box.data), ptr, ctypes.byref(box.dts),
                                 ctypes.byref(box.grid))
            assert rc == 0
            assert_states_equal(box.get_state(), ref, "AdvanceStep (%s) step %d" % (emf, s))
            assert box.dts.invDt_hyp == r["invdt"]
            dt = oracle_lib.next_dt(r["invdt"], 0.4, 1.1, dt)
    finally:
        assert abi.PGPU_SetEMFAverage(EMF_AVERAGES["uct_hll"]) == 0


@pytest.mark.parametrize("emf", AVERAGES)
def test_pluto_main_loop_with_library(native, emf):
    from oracle import refrun
    from pluto_amd import problems as P
    exe = os.path.join(refrun.REF_DIR, "pluto3d%s_dropin" % refrun.EMF_BUILDS[emf])
    if not os.path.exists(exe) or not refrun.available(3, emf):
        pytest.skip("drop-in executable not built (make -C oracle/ref dropin)")
    _say("case: PLUTO main loop with CT_EMF_AVERAGE %s linked against libpluto_amd.so, "
         "12^3, 6 steps" % emf)
    st = P.random_mhd_3d(12, seed=37)
    cpu = refrun.run(st, 6, dump=[1, 3], first_dt=1e-4, emf=emf)
    gpu = refrun.run(st, 6, dump=[1, 3], first_dt=1e-4, exe=exe)
    for s in cpu:
        (hc, xc), (hg, xg) = cpu[s], gpu[s]
        assert_states_equal(xg, xc, "main loop (%s) step %d" % (emf, s))
        assert hg["dt"] == hc["dt"] and hg["maxmach"] == hc["maxmach"]


def test_uct_hll_only_configurations_refuse(native, abi):
    from pluto_amd import problems as P
    _say("case: configurations that keep UCT_HLL data refuse the other averages")
    st = P.random_mhd_3d(8, seed=1)
    with pytest.raises(native.PGPUError):
        native.Domain.from_state(st, precision=native.PGPU_PREC_FP32_FLUX,
                                 emf_average=native.EMF_ARITHMETIC)
    with pytest.raises(native.PGPUError):
        native.Domain.from_state(st, pipeline=native.PIPELINE_MODULAR,
                                 emf_average=native.EMF_UCT0)
    dom = native.Domain.from_state(st, emf_average=native.EMF_UCT_CONTACT)
    from pluto_amd.native import check, lib
    with pytest.raises(native.PGPUError):
        check(lib().pgpu_set_precision(dom.h, native.PGPU_PREC_FP32_FLUX))
    with pytest.raises(native.PGPUError):
        check(lib().pgpu_set_pipeline(dom.h, native.PIPELINE_MODULAR))
    with pytest.raises(native.PGPUError):
        dom.set_emf_average(7)
    dom.set_emf_average("uct_hll")     # back to the default: FP32 allowed again
    check(lib().pgpu_set_precision(dom.h, native.PGPU_PREC_FP32_FLUX))
    dom.close()
    assert abi.PGPU_SetEMFAverage(0) == 1
    assert abi.PGPU_SetEMFAverage(5) == 1
