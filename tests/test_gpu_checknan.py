"""CheckNaN (tools.c:20-38): PLUTO tests every stage input for NaN and stops
the run with exit status 1 (update_stage.c:192, QUIT_PLUTO).

The native domain counts the stage inputs holding a NaN (the state it was
given, and the primitive state each stage produces for the next one) and
pgpu_sync reports PGPU_ENAN; the final state of a run, which no stage takes
as input, is not counted, as in PLUTO.  The library never exits: the
drop-in AdvanceStep reports the NaN at the step where PLUTO stops, as an
error that changes nothing, and the drop-in build's main loop (which checks
Integrate()'s return value) exits with status 1 at that step.  Checked against PLUTO's own main loop and the
same loop linked against the library, for a NaN in the initial state and
for a state whose first stage produces one (a velocity of 1e200 overflows
the energy).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


def _states():
    from pluto_amd import problems as P
    st = P.random_mhd_3d(8, seed=3)
    nan = st.copy()
    nan.prim[0, 3, 4, 5] = np.nan
    hot = st.copy()
    hot.prim[1, 3, 4, 5] = 1e200
    return st, nan, hot


@pytest.mark.parametrize("pipeline", [0, 1], ids=["fused", "modular"])
def test_domain_reports_nan_inputs(native, pipeline):
    _say("case: CheckNaN on the domain, pipeline %d" % pipeline)
    st, nan, hot = _states()
    dom = native.Domain.from_state(st, pipeline=pipeline)
    dom.set_dt(1e-3, 0.4, 1.1)
    dom.run(3)
    dom.sync()                                   # clean state: no report
    dom.close()
    for name, s in (("NaN in the initial state", nan), ("NaN from the first stage", hot)):
        dom = native.Domain.from_state(s, pipeline=pipeline)
        dom.set_dt(1e-3, 0.4, 1.1)
        dom.run(1)
        with pytest.raises(native.PGPUError, match="NaN"):
            dom.sync()
        dom.close()
    # a NaN state replaced by a clean one clears the report
    dom = native.Domain.from_state(nan, pipeline=pipeline)
    dom.set_state(st)
    dom.set_dt(1e-3, 0.4, 1.1)
    dom.run(2)
    dom.sync()
    dom.close()


def _run(exe, st, nsteps):
    from oracle import refrun
    wd = tempfile.mkdtemp(prefix="checknan_")
    with open(os.path.join(wd, "pluto.ini"), "w") as f:
        f.write(refrun._ini(st, 0.4, 1.1, 1e-3))
    refrun.write_state(os.path.join(wd, "state_in.bin"), st)
    env = dict(os.environ)
    env["PLUTO_REF_DUMP"] = "all"
    p = subprocess.run([exe, "-maxsteps", str(nsteps)], cwd=wd, env=env, capture_output=True,
                       text=True, timeout=120)
    dumps = {int(x[6:12]): refrun.read_state(os.path.join(wd, x), st)[1]
             for x in os.listdir(wd) if x.startswith("state_0")}
    return p.returncode, dumps, p.stdout + p.stderr


def test_pluto_main_loop_with_library_reports_nan(native):
    from conftest import assert_states_equal
    from oracle import refrun
    exe = os.path.join(refrun.REF_DIR, "pluto3d_dropin")
    if not os.path.exists(exe) or not refrun.available(3):
        pytest.skip("drop-in executable not built (make -C oracle/ref dropin)")
    st, nan, hot = _states()
    _say("case: PLUTO main loop vs drop-in, clean state")
    rc_r, ref, _ = _run(refrun.exe_path(3), st, 3)
    rc_g, gpu, out = _run(exe, st, 3)
    assert rc_r == rc_g == 0 and sorted(ref) == sorted(gpu)
    assert "NaN found" not in out
    for k in ref:
        assert_states_equal(gpu[k], ref[k], "clean run, step %d" % k)
    for name, s in (("NaN initial state", nan), ("NaN from stage 1", hot)):
        _say("case: PLUTO main loop vs drop-in, %s" % name)
        rc_r, ref, ref_out = _run(refrun.exe_path(3), s, 3)
        assert rc_r == 1 and sorted(ref) == [0] and "NaN found" in ref_out
        rc_g, gpu, out = _run(exe, s, 3)
        # the library reports the NaN at the same step and leaves the data as
        # it was; the drop-in's main loop checks Integrate()'s return value
        # and stops there with PLUTO's exit status
        assert "NaN found in a stage input" in out, out[-2000:]
        assert rc_g == 1, (name, rc_g, out[-2000:])
        assert sorted(gpu) == sorted(ref) == [0], name
        assert _same(gpu[0], ref[0]), name + ", step 0"


def _same(a, b):
    """Bitwise-equal states, NaN equal to NaN."""
    return all(np.array_equal(x, y, equal_nan=True)
               for x, y in ((a.prim, b.prim), (a.bx, b.bx), (a.by, b.by), (a.bz, b.bz)))
