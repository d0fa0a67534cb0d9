"""NextTimeStep's stop (main.c:534-543): PLUTO quits with exit status 1 when
the next dt falls below 1e-9 of first_dt.  The device dt control records it
in its step-end kernel, takes no more time in the steps enqueued after it,
and pgpu_sync / pgpu_batch_sync report PGPU_EDTSMALL.

The condition is forced from the host with a tiny CFL number (the next dt
is CFL x the CFL limit, far below 1e-9 first_dt) on an ordinary state, so
every kernel runs on sane data and exits normally.
"""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import assert_states_equal

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


def _reference(st, cfl, first_dt, nsteps):
    from oracle import refrun
    wd = tempfile.mkdtemp(prefix="dtsmall_")
    with open(os.path.join(wd, "pluto.ini"), "w") as f:
        f.write(refrun._ini(st, cfl, 1.1, first_dt))
    refrun.write_state(os.path.join(wd, "state_in.bin"), st)
    env = dict(os.environ)
    env["PLUTO_REF_DUMP"] = "all"
    p = subprocess.run([refrun.exe_path(3), "-maxsteps", str(nsteps)], cwd=wd, env=env,
                       capture_output=True, text=True, timeout=120)
    dumps = {int(x[6:12]): refrun.read_state(os.path.join(wd, x), st)
             for x in os.listdir(wd) if x.startswith("state_0")}
    return p.returncode, dumps, p.stdout + p.stderr


def test_domain_stops_where_pluto_quits(native, oracle_lib):
    from oracle import refrun
    from pluto_amd import problems as P
    if not refrun.available(3):
        pytest.skip("reference executables not built (make -C oracle/ref)")
    st = P.random_mhd_3d(8, seed=17)
    cfl, first_dt = 1e-12, 1e-3
    _say("case: dt too small, 8^3, CFL %g, first_dt %g" % (cfl, first_dt))
    rc, dumps, out = _reference(st, cfl, first_dt, 4)
    # PLUTO takes step 0 -> 1, then NextTimeStep quits before step 1's dump
    assert rc == 1 and sorted(dumps) == [0], out[-2000:]
    m = re.search(r"dt is too small \(\s*([0-9.e+-]+)\)", out)
    assert m, out[-2000:]
    dom = native.Domain.from_state(st)
    dom.set_dt(first_dt, cfl, 1.1)
    dom.run(4)
    with pytest.raises(native.PGPUError, match="dt is too small") as e:
        dom.sync()
    # the same next dt as PLUTO's, in PLUTO's print format
    assert m.group(1) in str(e.value), (m.group(1), str(e.value))
    sc = dom.scalars()
    # one step of first_dt was taken; the later ones took dt = 0
    assert sc["time"] == first_dt and sc["step"] == 4 and sc["dt"] == 0.0
    # the state of the step whose next dt was too small: one step of the
    # oracle (pinned to PLUTO's main loop)
    dom2 = native.Domain.from_state(st)
    dom2.set_dt(first_dt, cfl, 1.1)
    dom2.run(1)
    with pytest.raises(native.PGPUError, match="dt is too small"):
        dom2.sync()
    ref = st.copy()
    oracle_lib.rk2_step(ref, first_dt)
    assert_states_equal(dom2.get_state(st), ref, "step 1")
    # a reasonable CFL runs on without a report; set_dt clears it
    dom2.set_state(st)
    dom2.set_dt(first_dt, 0.4, 1.1)
    dom2.run(3)
    dom2.sync()
    dom.close()
    dom2.close()


def test_batch_reports_dt_too_small(native):
    from pluto_amd import problems as P
    _say("case: dt too small in a patch batch")
    edges = [8, 12, 8]
    b = native.Batch(edges)
    for p, n in enumerate(edges):
        b.set_patch(p, P.random_mhd_3d(n, seed=30 + p))
    b.set_dt(1e-3, 0.4, 1.1)
    b.run(2)
    b.sync()
    b.set_dt(1e-3, 1e-12, 1.1)
    b.run(2)
    with pytest.raises(native.PGPUError, match="dt is too small"):
        b.sync()
    for p in range(len(edges)):
        assert b.scalars(p)["time"] == 1e-3
    b.close()
