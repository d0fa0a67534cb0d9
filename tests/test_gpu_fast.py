"""Tolerance-mode FP64 flux (PGPU_PREC_FP64_FAST): the FP64 flux kernels
compiled with FMA contraction and Newton-refined division and square root.

Bar: the conserved variables against the oracle (PLUTO's FP64 CPU path, bit
for bit) within 1e-12 relative per step and 1e-9 after 100 steps (L-infinity
per field over its largest magnitude), div.B at round-off (the CT update is
the bit-exact FP64 one), and switching back to PGPU_PREC_FP64 bit-exact
again.
"""
import numpy as np
import pytest

from conftest import assert_states_equal

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "mx", "my", "mz", "E", "bx", "by", "bz")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


def _cons(st, gamma=5.0 / 3.0):
    v = st.prim
    kin = v[1] ** 2 + v[2] ** 2 + v[3] ** 2
    e = 0.5 * (v[0] * kin + v[4] ** 2 + v[5] ** 2 + v[6] ** 2) + v[7] / (gamma - 1.0)
    return [v[0], v[0] * v[1], v[0] * v[2], v[0] * v[3], e, st.bx, st.by, st.bz]


def _rel_linf(ref, st):
    """max |difference| over max |reference| per field (the difference itself
    where the reference field is identically zero, e.g. m_z and B_z in 2D)."""
    out = {}
    for n, a, b in zip(FIELDS, _cons(ref), _cons(st)):
        scale = np.abs(a).max()
        out[n] = float(np.abs(b - a).max() / (scale if scale > 0 else 1.0))
    return out


def _fast(native, st, **kw):
    return native.Domain.from_state(st, precision=native.PGPU_PREC_FP64_FAST, **kw)


def test_one_step_within_1e12(native, oracle_lib):
    from pluto_amd import problems as P
    _say("case: FP64 fast, one step, random 48^3")
    st = P.random_mhd_3d(48, seed=2024)
    ref = st.copy()
    oracle_lib.rk2_step(ref, 2e-4)
    dom = _fast(native, st)
    dom.set_dt(2e-4, fixed=True)
    dom.run(1)
    out = dom.get_state(st)
    err = _rel_linf(ref, out)
    _say("  relative L-inf per field: " + ", ".join("%s %.1e" % kv for kv in err.items()))
    assert max(err.values()) <= 1e-12
    assert max(err.values()) > 0.0          # not the bit-exact kernels
    dom.close()


def test_100_steps_within_1e9(native, oracle_lib):
    from pluto_amd import problems as P
    _say("case: FP64 fast, 100 CFL steps, random 48^3")
    st = P.random_mhd_3d(48, seed=7)
    ref = st.copy()
    dts, _ = oracle_lib.run(ref, 100, 1e-4, 0.4, 1.1)
    dom = _fast(native, st)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(100)
    out = dom.get_state(st)
    gdt, _, _ = dom.history(100)
    assert np.allclose(gdt, dts, rtol=1e-11, atol=0)
    err = _rel_linf(ref, out)
    _say("  relative L-inf per field: " + ", ".join("%s %.1e" % kv for kv in err.items()))
    assert max(err.values()) <= 1e-9
    divb = np.abs(P.divb(out)).max()
    _say("  max |div B| %.1e" % divb)
    assert divb <= 1e-11
    dom.close()


def test_orszag_tang_2d_200_steps(native, oracle_lib):
    from pluto_amd import problems as P
    _say("case: FP64 fast, Orszag-Tang 128^2, 200 CFL steps")
    st = P.orszag_tang_2d(128)
    ref = st.copy()
    oracle_lib.run(ref, 200, 1e-4, 0.4, 1.1)
    dom = _fast(native, st)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(200)
    err = _rel_linf(ref, dom.get_state(st))
    _say("  relative L-inf per field: " + ", ".join("%s %.1e" % kv for kv in err.items()))
    assert max(err.values()) <= 1e-9
    dom.close()


def test_switch_back_is_bit_exact_and_refusals(native, oracle_lib):
    from pluto_amd import problems as P
    from pluto_amd.native import lib, check
    st = P.random_mhd_3d(12, seed=9)
    ref = st.copy()
    oracle_lib.rk2_step(ref, 1e-3)
    dom = _fast(native, st)
    check(lib().pgpu_set_precision(dom.h, native.PGPU_PREC_FP64))
    dom.set_dt(1e-3, fixed=True)
    dom.run(1)
    assert_states_equal(dom.get_state(st), ref, "FP64 after the fast mode")
    dom.set_solver(native.SOLVER_HLLC)
    with pytest.raises(native.PGPUError):
        check(lib().pgpu_set_precision(dom.h, native.PGPU_PREC_FP64_FAST))
    dom.close()
    with pytest.raises(native.PGPUError):
        native.Domain.from_state(st, pipeline=native.PIPELINE_MODULAR,
                                 precision=native.PGPU_PREC_FP64_FAST)
