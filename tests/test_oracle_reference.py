"""The oracle against the live reference build (oracle/_ref), larger cases.

Skipped where the reference has not been built (make -C oracle/ref).  The
reference runs PLUTO's own main loop; states, dt sequence and max Mach must
agree bit for bit.
"""
import numpy as np
import pytest

from conftest import assert_states_equal
from oracle import refrun
from pluto_amd import problems as P

pytestmark = pytest.mark.skipif(not refrun.available(3) or not refrun.available(2),
                                reason="reference not built (make -C oracle/ref)")


def _compare(oracle_lib, st, nsteps, dump, **kw):
    ref = refrun.run(st, nsteps, dump=dump, **kw)
    o = st.copy()
    if "fixed_dt" in kw:
        dt, cfl, var = kw["fixed_dt"], 1e30, 1.0
    else:
        dt, cfl, var = kw.get("first_dt", 1e-4), kw.get("cfl", 0.4), kw.get("cfl_max_var", 1.1)
    done = 0
    for s in sorted(ref):
        while done < s:
            r = oracle_lib.rk2_step(o, dt)
            dt = oracle_lib.next_dt(r["invdt"], cfl, var, dt)
            done += 1
        h, x = ref[s]
        assert_states_equal(o, x, "step %d" % s)
        assert dt == h["dt"]
        assert r["maxmach"] == h["maxmach"]


def test_random_3d_cfl(oracle_lib):
    _compare(oracle_lib, P.random_mhd_3d(16, seed=77), 12, [1, 6], first_dt=1e-4)


def test_orszag_tang_2d(oracle_lib):
    _compare(oracle_lib, P.orszag_tang_2d(64), 60, [20, 40], first_dt=1e-4)


def test_blast_3d_strong_shocks(oracle_lib):
    # pressure ratio 1000: supersonic outflow, every HLLD region, large dt changes
    _compare(oracle_lib, P.blast_3d(16), 10, [3], first_dt=1e-5)


def test_smooth_adjoint_state(oracle_lib):
    _compare(oracle_lib, P.smooth_mhd_3d(12, seed=5), 5, [], fixed_dt=2e-3)
