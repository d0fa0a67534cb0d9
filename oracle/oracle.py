"""ctypes wrapper of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the CPU restatement the HIP path is checked
against.  Imported by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py -- never by the product package.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ora_rk2_step.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 4 + [_dp] * 4 + [
            ctypes.c_double, _dp, _dp, ctypes.POINTER(ctypes.c_ubyte), _dp, _dp, _dp]
        L.ora_rk2_step.restype = ctypes.c_int
        L.ora_next_dt.argtypes = [ctypes.c_double] * 4
        L.ora_next_dt.restype = ctypes.c_double
        L.ora_run.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 4 + [_dp] * 4 + [
            ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
            _dp, _dp]
        L.ora_run.restype = ctypes.c_int
        L.ora_plm_row.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.ora_hlld_rows.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                    _dp, _dp, _dp, _dp, _dp, _dp]
        L.ora_p2c_rows.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp]
        L.ora_c2p_rows.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp, _dp,
                                   ctypes.POINTER(ctypes.c_ubyte)]
        L.ora_c2p_rows.restype = ctypes.c_int
        L.ora_num_threads.restype = ctypes.c_int
        L.ora_set_threads.argtypes = [ctypes.c_int]
        L.ora_set_solver.argtypes = [ctypes.c_int]
        L.ora_set_emf.argtypes = [ctypes.c_int]
        L.ora_set_emf.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def num_threads():
    return lib().ora_num_threads()


def set_threads(n):
    lib().ora_set_threads(int(n))


SOLVERS = {"hlld": 0, "hll": 1, "hllc": 2, "tvdlf": 3, "roe": 4}


# CT_EMF_AVERAGE labels (ct.h:18-21)
EMF_AVERAGES = {"arithmetic": 1, "uct0": 2, "uct_contact": 3, "uct_hll": 4}


def _solver(solver, emf="uct_hll"):
    lib().ora_set_solver(SOLVERS[solver])
    if lib().ora_set_emf(EMF_AVERAGES[emf]) != 0:
        raise ValueError(emf)


def rk2_step(st, dt, gamma=5.0 / 3.0, want_emf=False, solver="hlld", emf="uct_hll"):
    """One RK2 step of ``st`` (MHDState) in place with the Riemann solver
    ``solver`` (a key of SOLVERS) and the EMF average ``emf`` (a key of
    EMF_AVERAGES).

    Returns dict(invdt, maxmach, flag, fail[, ex, ey, ez of stage 1])."""
    _solver(solver, emf)
    for a in (st.prim, st.bx, st.by, st.bz):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    nz, ny, nx = st.shape
    dx, dy, dz = st.spacing()
    inv = np.zeros(1)
    mach = np.zeros(1)
    flag = np.zeros(st.shape, dtype=np.uint8)
    emf = [np.zeros(st.shape) for _ in range(3)] if want_emf else [None] * 3
    rc = lib().ora_rk2_step(nx, ny, nz, st.dims, dx, dy, dz, gamma,
                            _p(st.prim), _p(st.bx), _p(st.by), _p(st.bz), dt,
                            _p(inv), _p(mach), flag.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                            _p(emf[0]), _p(emf[1]), _p(emf[2]))
    if rc < 0:
        raise ValueError("oracle rejected the box")
    out = dict(invdt=float(inv[0]), maxmach=float(mach[0]), flag=flag, fail=rc)
    if want_emf:
        out.update(ex=emf[0], ey=emf[1], ez=emf[2])
    return out


def next_dt(invdt, cfl, cfl_max_var, dt_old):
    return lib().ora_next_dt(invdt, cfl, cfl_max_var, dt_old)


def run(st, nsteps, dt0, cfl=0.4, cfl_max_var=1.1, fixed_dt=False, gamma=5.0 / 3.0,
        solver="hlld", emf="uct_hll"):
    """nsteps RK2 steps in place with the reference's dt control.
    Returns (dts, maxmach) arrays of length nsteps."""
    _solver(solver, emf)
    nz, ny, nx = st.shape
    dx, dy, dz = st.spacing()
    dts = np.zeros(nsteps)
    mach = np.zeros(nsteps)
    lib().ora_run(nx, ny, nz, st.dims, dx, dy, dz, gamma, _p(st.prim), _p(st.bx),
                  _p(st.by), _p(st.bz), nsteps, dt0, cfl, cfl_max_var, int(bool(fixed_dt)),
                  _p(dts), _p(mach))
    return dts, mach


def plm_row(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    n = v.shape[0]
    vp = np.full((n, 8), np.nan)
    vm = np.full((n, 8), np.nan)
    lib().ora_plm_row(n, _p(v), _p(vp), _p(vm))
    return vp, vm


def hlld_rows(d, vl, vr, gamma=5.0 / 3.0, solver="hlld"):
    _solver(solver)
    vl = np.ascontiguousarray(vl, dtype=np.float64)
    vr = np.ascontiguousarray(vr, dtype=np.float64)
    n = vl.shape[0]
    flux = np.zeros((n, 8))
    press, sl, sr, cmax = (np.zeros(n) for _ in range(4))
    mach = np.zeros(1)
    lib().ora_hlld_rows(d, n, _p(vl), _p(vr), gamma, _p(flux), _p(press), _p(sl),
                        _p(sr), _p(cmax), _p(mach))
    return flux, press, sl, sr, cmax, float(mach[0])


def p2c_rows(v, gamma=5.0 / 3.0):
    v = np.ascontiguousarray(v, dtype=np.float64)
    u = np.zeros_like(v)
    lib().ora_p2c_rows(v.shape[0], _p(v), gamma, _p(u))
    return u


def c2p_rows(u, gamma=5.0 / 3.0):
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.zeros_like(u)
    uo = np.zeros_like(u)
    flag = np.zeros(u.shape[0], dtype=np.uint8)
    fail = lib().ora_c2p_rows(u.shape[0], _p(u), gamma, _p(v), _p(uo),
                              flag.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    return v, uo, flag, fail
