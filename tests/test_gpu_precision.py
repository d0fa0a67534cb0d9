"""FP32-flux precision path: PLM + HLLD in FP32, everything else FP64.

Bar: relative L1 error per field <= 1e-6 against the FP64 result (the FP64
GPU path, itself bit-identical to the oracle, and the oracle directly), and
div.B at round-off because the CT update stays FP64.  The device-side error
reduction used to report these figures is checked against numpy first.
"""
import numpy as np
import pytest

from conftest import assert_states_equal

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "mx", "my", "mz", "E", "bx", "by", "bz")
L1_TOL = 1e-6        # relative L1 per field, FP32 flux vs FP64


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


def _fields(st, gamma=5.0 / 3.0):
    v = st.prim
    kin = v[1] ** 2 + v[2] ** 2 + v[3] ** 2
    e = 0.5 * (v[0] * kin + v[4] ** 2 + v[5] ** 2 + v[6] ** 2) + v[7] / (gamma - 1.0)
    return [v[0], v[0] * v[1], v[0] * v[2], v[0] * v[3], e, st.bx, st.by, st.bz]


def _np_error(ref, st):
    out = {}
    for name, a, b in zip(FIELDS, _fields(ref), _fields(st)):
        d = np.abs(b - a)
        out[name] = (d.sum() / np.abs(a).sum(), d.max() / np.abs(a).max())
    return out


def test_state_error_matches_numpy(native):
    from pluto_amd import problems as P
    _say("case: pgpu_state_error vs numpy, 16^3, dt 1e-3 vs 1.1e-3")
    st = P.random_mhd_3d(16, seed=21)
    a = native.Domain.from_state(st)
    b = native.Domain.from_state(st)
    a.set_dt(1e-3, fixed=True)
    b.set_dt(1.1e-3, fixed=True)
    a.run(1)
    b.run(1)
    dev = b.error_vs(a)
    host = _np_error(a.get_state(st), b.get_state(st))
    for name in FIELDS:
        for x, y in zip(dev[name], host[name]):
            assert y > 0
            assert abs(x - y) <= 1e-12 * y, (name, x, y)
    same = a.error_vs(a)
    assert all(v == (0.0, 0.0) for v in same.values())


def test_fp32_flux_error_budget(native):
    """32^3 random field, 20 CFL-sized fixed steps: FP32-flux vs FP64."""
    from pluto_amd import problems as P
    _say("case: FP32 flux vs FP64, 32^3, 20 steps")
    st = P.random_mhd_3d(32, seed=7)
    ref = native.Domain.from_state(st)
    f32 = native.Domain.from_state(st, precision=native.PGPU_PREC_FP32_FLUX)
    for d in (ref, f32):
        d.set_dt(2e-3, fixed=True)
        d.run(20)
    err = f32.error_vs(ref)
    _say("  per-field (L1_rel, Linf_rel): %s" % {k: ("%.2e" % a, "%.2e" % b)
                                                for k, (a, b) in err.items()})
    for name, (l1, linf) in err.items():
        assert l1 <= L1_TOL, (name, l1)
        assert linf <= 100 * L1_TOL, (name, linf)
    # the FP32 path must really run in FP32
    assert err["rho"][0] > 1e-12
    out = f32.get_state(st)
    db = P.divb(out)
    bmax = max(np.abs(out.bx).max(), np.abs(out.by).max(), np.abs(out.bz).max())
    assert np.abs(db).max() * out.spacing()[0] / bmax <= 1e-13


def test_fp32_flux_2d_orszag_tang(native):
    """2D boxes take the per-cell EMF kernel over the float face data."""
    from pluto_amd import problems as P
    _say("case: FP32 flux vs FP64, Orszag-Tang 64^2, 40 steps")
    st = P.orszag_tang_2d(64)
    ref = native.Domain.from_state(st)
    f32 = native.Domain.from_state(st, precision=native.PGPU_PREC_FP32_FLUX)
    for d in (ref, f32):
        d.set_dt(1e-3, fixed=True)     # Courant number about 0.3
        d.run(40)
    err = f32.error_vs(ref)
    _say("  per-field (L1_rel, Linf_rel): %s" % {k: ("%.2e" % a, "%.2e" % b)
                                                for k, (a, b) in err.items()})
    for name in ("rho", "mx", "my", "E", "bx", "by"):
        l1, linf = err[name]
        assert l1 <= L1_TOL, (name, l1)
        assert linf <= 100 * L1_TOL, (name, linf)
    assert err["rho"][0] > 1e-12
    out = f32.get_state(st)
    db = P.divb(out)
    bmax = max(np.abs(out.bx).max(), np.abs(out.by).max())
    assert np.abs(db).max() * out.spacing()[0] / bmax <= 1e-13


def test_fp32_flux_vs_oracle(native, oracle_lib):
    from pluto_amd import problems as P
    _say("case: FP32 flux vs oracle, 16^3, CFL steps")
    st = P.random_mhd_3d(16, seed=8)
    ref = st.copy()
    dts, _ = oracle_lib.run(ref, 8, 1e-4, 0.4, 1.1)
    dom = native.Domain.from_state(st, precision=native.PGPU_PREC_FP32_FLUX)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(8)
    gdt, _, _ = dom.history(8)
    # dt follows the FP32 signal speeds: equal to about 1e-7 relative
    assert np.allclose(gdt, dts, rtol=1e-6, atol=0)
    out = dom.get_state(st)
    for name, (l1, linf) in _np_error(ref, out).items():
        assert l1 <= L1_TOL, (name, l1)


def test_fp64_unchanged_by_precision_switch(native, oracle_lib):
    """Switching to FP32 and back leaves the FP64 path bit-exact."""
    from pluto_amd import problems as P
    _say("case: precision switch round trip, 12^3")
    st = P.random_mhd_3d(12, seed=9)
    ref = st.copy()
    oracle_lib.rk2_step(ref, 1e-3)
    dom = native.Domain.from_state(st, precision=native.PGPU_PREC_FP32_FLUX)
    from pluto_amd.native import lib, check
    check(lib().pgpu_set_precision(dom.h, native.PGPU_PREC_FP64))
    dom.set_dt(1e-3, fixed=True)
    dom.run(1)
    assert_states_equal(dom.get_state(st), ref, "FP64 after switch")


def test_fp32_needs_fused_pipeline(native):
    from pluto_amd import problems as P
    st = P.random_mhd_3d(8, seed=1)
    with pytest.raises(native.PGPUError):
        native.Domain.from_state(st, pipeline=native.PIPELINE_MODULAR,
                                 precision=native.PGPU_PREC_FP32_FLUX)
