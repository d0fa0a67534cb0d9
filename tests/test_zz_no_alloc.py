"""pgpu_riemann_rows_static, the Riemann row entry point that allocates
nothing, against the allocating row entry points (HLLD_Solver, HLL_Solver,
HLLC_Solver, LF_Solver, Roe_Solver) on the reference's golden rows: bit for
bit, on two calls in a row (the second with fewer rows, so stale scratch
from the first would show), and a refusal above the fixed size.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, golden
from pluto_amd import plutoabi as A

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "pluto_amd", "libpluto_amd.so")
SOLVERS = (("hlld", 0, "HLLD_Solver"), ("hll", 1, "HLL_Solver"), ("hllc", 2, "HLLC_Solver"),
           ("tvdlf", 3, "LF_Solver"), ("roe", 4, "Roe_Solver"))


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def abi():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    lib = A.load_abi(LIB)
    f = lib.pgpu_riemann_rows_static
    f.argtypes = [ctypes.c_int, ctypes.POINTER(A.Sweep), ctypes.c_int, ctypes.c_int,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    return lib


def _rows(name, d, lo, hi):
    z = golden("rows_%s.npz" % name)
    return z["%s%d_vl" % (name, d)][lo:hi], z["%s%d_vr" % (name, d)][lo:hi]


def _call(abi, fn, vl, vr, d):
    """One call over rows 0..n-1 of a fresh sweep; returns its outputs."""
    n = vl.shape[0]
    S = A.SweepRows(n)
    S.vL[:n] = vl
    S.vR[1:n + 1] = vr
    S.flux[:] = np.nan
    cmax = np.full(n, np.nan)
    ctypes.c_int.in_dll(abi, "g_dir").value = d
    ctypes.c_double.in_dll(abi, "g_maxMach").value = 0.0
    rc = fn(S, 0, n - 1, cmax)
    mach = ctypes.c_double.in_dll(abi, "g_maxMach").value
    return rc, (S.flux[:n].copy(), S.press[:n].copy(), S.SL[:n].copy(), S.SR[:n].copy(),
                cmax, mach)


@pytest.mark.parametrize("name,code,ref_fn", SOLVERS)
def test_static_rows_equal_allocating_rows(abi, name, code, ref_fn):
    ref = getattr(abi, ref_fn)
    for d in (0, 1, 2):
        # a full set of rows, then a smaller one from another part of the table
        for lo, hi in ((0, 600), (100, 137)):
            _say("case: %s static rows dir %d rows %d..%d" % (name, d, lo, hi - 1))
            vl, vr = _rows(name, d, lo, hi)
            rc0, want = _call(abi, lambda S, b, e, c: ref(ctypes.byref(S.sw), b, e,
                                                          c.ctypes.data, None), vl, vr, d)
            rc1, got = _call(abi, lambda S, b, e, c: abi.pgpu_riemann_rows_static(
                code, ctypes.byref(S.sw), b, e, c.ctypes.data), vl, vr, d)
            assert rc1 == 0
            for a, b in zip(want[:5], got[:5]):
                assert np.array_equal(a, b)
            assert want[5] == got[5]


def test_static_rows_refuse_above_limit(abi):
    _say("case: static rows above the fixed size")
    nmax = 8192
    n = nmax + 1
    S = A.SweepRows(n)
    S.vL[:n] = [1.0, 0.1, 0.0, 0.0, 0.5, 0.2, 0.1, 1.0]
    S.vR[1:n + 1] = S.vL[:n]
    S.flux[:] = -7.0
    cmax = np.full(n, -7.0)
    ctypes.c_int.in_dll(abi, "g_dir").value = 0
    rc = abi.pgpu_riemann_rows_static(0, ctypes.byref(S.sw), 0, n - 1, cmax.ctypes.data)
    assert rc == -1
    assert (S.flux[:n] == -7.0).all() and (cmax == -7.0).all()
    # exactly the maximum is served, and equals HLLD_Solver
    rc = abi.pgpu_riemann_rows_static(0, ctypes.byref(S.sw), 0, nmax - 1, cmax.ctypes.data)
    assert rc == 0
    f1 = S.flux[:nmax].copy()
    c1 = cmax[:nmax].copy()
    abi.HLLD_Solver(ctypes.byref(S.sw), 0, nmax - 1, cmax.ctypes.data, None)
    assert np.array_equal(S.flux[:nmax], f1) and np.array_equal(cmax[:nmax], c1)
    assert abi.pgpu_riemann_rows_static(9, ctypes.byref(S.sw), 0, 3, cmax.ctypes.data) == -1
