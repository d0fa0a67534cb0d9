"""The oracle against the reference's golden vectors (tests/golden/*.npz),
for the HLLD, HLL and HLLC Riemann solvers.

The vectors were produced by the reference itself (PLUTO's row kernels and
its main loop, built from its sources; tests/golden/make_fixtures.py).  The
oracle must reproduce them bit for bit.
"""
import numpy as np
import pytest

from conftest import assert_states_equal, golden, load_run


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hlld_rows(oracle_lib, d):
    z = golden("rows_hlld.npz")
    flux, press, sl, sr, cmax, mach = oracle_lib.hlld_rows(d, z["hlld%d_vl" % d], z["hlld%d_vr" % d])
    assert np.array_equal(flux, z["hlld%d_flux" % d])
    assert np.array_equal(press, z["hlld%d_press" % d])
    assert np.array_equal(sl, z["hlld%d_sl" % d])
    assert np.array_equal(sr, z["hlld%d_sr" % d])
    assert np.array_equal(cmax, z["hlld%d_cmax" % d])
    assert mach == z["hlld%d_mach" % d][0]


@pytest.mark.parametrize("solver", ["hll", "hllc", "tvdlf", "roe"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_solver_rows(oracle_lib, solver, d):
    z = golden("rows_%s.npz" % solver)
    k = solver
    flux, press, sl, sr, cmax, mach = oracle_lib.hlld_rows(d, z["%s%d_vl" % (k, d)],
                                                           z["%s%d_vr" % (k, d)], solver=solver)
    assert np.array_equal(flux, z["%s%d_flux" % (k, d)])
    assert np.array_equal(press, z["%s%d_press" % (k, d)])
    assert np.array_equal(sl, z["%s%d_sl" % (k, d)])
    assert np.array_equal(sr, z["%s%d_sr" % (k, d)])
    assert np.array_equal(cmax, z["%s%d_cmax" % (k, d)])
    assert mach == z["%s%d_mach" % (k, d)][0]
    # the fixture is this solver's, not HLLD's: intermediate-state fluxes differ
    hflux = oracle_lib.hlld_rows(d, z["%s%d_vl" % (k, d)], z["%s%d_vr" % (k, d)])[0]
    assert np.sum(hflux != flux) > flux.size // 4


@pytest.mark.parametrize("d", [0, 1, 2])
def test_plm_rows(oracle_lib, d):
    z = golden("rows_plm.npz")
    v, bn = z["plm%d_v" % d], z["plm%d_bn" % d]
    vp, vm = oracle_lib.plm_row(v)
    rvp, rvm = z["plm%d_vp" % d], z["plm%d_vm" % d]
    n = v.shape[0]
    # the reference overwrites the normal field with the staggered value:
    # vp[i][BXn] = bn[i], vm[i][BXn] = bn[i-1] (plm_states.c:271-274)
    vp[1:n - 1, 4 + d] = bn[1:n - 1]
    vm[1:n - 1, 4 + d] = bn[0:n - 2]
    assert np.array_equal(vp[1:n - 1], rvp[1:n - 1])
    assert np.array_equal(vm[1:n - 1], rvm[1:n - 1])
    assert np.array_equal(oracle_lib.p2c_rows(vp[1:n - 1]), z["plm%d_up" % d][1:n - 1])
    assert np.array_equal(oracle_lib.p2c_rows(vm[1:n - 1]), z["plm%d_um" % d][1:n - 1])


def test_conversions(oracle_lib):
    z = golden("rows_conv.npz")
    assert np.array_equal(oracle_lib.p2c_rows(z["p2c_v"]), z["p2c_u"])
    v, uo, flag, fail = oracle_lib.c2p_rows(z["c2p_u"])
    assert np.array_equal(v, z["c2p_v"])
    assert np.array_equal(uo, z["c2p_uout"])
    assert np.array_equal(flag, z["c2p_flag"])
    assert fail == 1 and int(np.sum(flag == 64)) == 60


EMF_AVERAGES = ["arithmetic", "uct0", "uct_contact"]
EMF_RUNS = ["step_%s_%s.npz" % (box, emf) for emf in EMF_AVERAGES
            for box in ("3d_8", "blob_12x10x8", "ot2d_32")]


@pytest.mark.parametrize("name", ["step_3d_8_fixed.npz", "step_3d_8_cfl.npz",
                                  "step_3d_14x12x10_cfl.npz", "step_ot2d_32.npz",
                                  "step_3d_8_hll.npz", "step_ot2d_32_hll.npz",
                                  "step_3d_8_hllc.npz", "step_ot2d_32_hllc.npz",
                                  "step_3d_8_tvdlf.npz", "step_ot2d_32_tvdlf.npz",
                                  "step_3d_8_roe.npz", "step_ot2d_32_roe.npz"] + EMF_RUNS)
def test_steps(oracle_lib, name):
    init, steps, run = load_run(name)
    solver = run.get("solver", "hlld")
    emf = run.get("emf", "uct_hll")
    st = init.copy()
    fixed = run.get("fixed_dt")
    if fixed is not None:
        dt0, cfl, var = fixed, 1.0e30, 1.0
    else:
        dt0, cfl, var = run["first_dt"], run["cfl"], run["cfl_max_var"]
    done = 0
    dt = dt0
    for s in sorted(steps):
        while done < s:
            r = oracle_lib.rk2_step(st, dt, solver=solver, emf=emf)
            dt = oracle_lib.next_dt(r["invdt"], cfl, var, dt)
            done += 1
        h, ref = steps[s]
        assert_states_equal(st, ref, "%s step %d" % (name, s))
        assert dt == h["dt"], "next dt after step %d" % s
        assert r["maxmach"] == h["maxmach"]


@pytest.mark.parametrize("solver", ["hll", "hllc", "tvdlf", "roe"])
def test_solver_run_is_not_hlld(oracle_lib, solver):
    """The reference honoured ``Solver <solver>``: its run differs from HLLD's."""
    init, steps, run = load_run("step_3d_8_%s.npz" % solver)
    st = init.copy()
    oracle_lib.run(st, 1, run["first_dt"], run["cfl"], run["cfl_max_var"], solver="hlld")
    assert np.sum(st.prim != steps[1][1].prim) > st.prim.size // 2


@pytest.mark.parametrize("emf", EMF_AVERAGES)
def test_emf_average_runs_differ(oracle_lib, emf):
    """The reference honoured CT_EMF_AVERAGE: one step of its run differs in
    the faces from UCT_HLL's and from the other averages'."""
    init, steps, run = load_run("step_3d_8_%s.npz" % emf)
    for other in ["uct_hll"] + [e for e in EMF_AVERAGES if e != emf]:
        st = init.copy()
        oracle_lib.run(st, 1, run["first_dt"], run["cfl"], run["cfl_max_var"], emf=other)
        assert np.sum(st.bx != steps[1][1].bx) > st.bx.size // 2, other


def test_uct_contact_takes_both_sign_branches(oracle_lib):
    """The blob state has faces with zero mass flux and faces with either
    sign in its first step, so the corner integration's three upwind cases
    are all exercised by its fixture."""
    init, steps, run = load_run("step_blob_12x10x8_uct_contact.npz")
    st = init.copy()
    r = oracle_lib.rk2_step(st, run["first_dt"], emf="uct_contact")
    v = init.prim[1:4]
    moving = np.abs(v).max(axis=0) > 0
    assert 0 < moving.sum() < moving.size // 2
