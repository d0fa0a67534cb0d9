"""Algorithmic floating-point operation count of one zone-update.

TEST AND MEASUREMENT INFRASTRUCTURE ONLY.  oracle/libopcount.so is the
oracle (mhd_oracle.c, the CPU restatement of PLUTO's step) compiled with
every double replaced by a counting type (opcount.hpp); it computes the same
IEEE values as liboracle.so.  One RK2 step of a periodic n^3 box costs

    ops(n) = A n^3 + B n^2      (PLM on n + 2 cells and n + 1 Riemann problems
                                 per row of n cells, as PLUTO's row ranges)

so two box sizes give A, the operations per zone-update of an unbounded box:
A = 2 ops(2m)/(2m)^3 - ops(m)/m^3.

    python -m oracle.opcount [--out profiles/flop_count_step.json]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libopcount.so")
KINDS = ("add", "mul", "div", "sqrt", "neg", "fabs", "cmp")

_lib = None


def lib():
    global _lib
    if _lib is None:
        from oracle import oracle as O
        if not os.path.exists(LIB_PATH):
            O.build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.ora_rk2_step.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 4 + [dp] * 4 + [
            ctypes.c_double, dp, dp, ctypes.POINTER(ctypes.c_ubyte), dp, dp, dp]
        L.ora_rk2_step.restype = ctypes.c_int
        L.ora_opcounts_get.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
        L.ora_set_solver.argtypes = [ctypes.c_int]
        L.ora_set_emf.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def counted_step(st, dt, gamma=5.0 / 3.0):
    """One RK2 step of st in place (HLLD, UCT_HLL) with the counting build;
    returns {kind: count}."""
    L = lib()
    L.ora_set_solver(0)
    L.ora_set_emf(4)
    dp = ctypes.POINTER(ctypes.c_double)
    nz, ny, nx = st.shape
    dx, dy, dz = st.spacing()
    inv, mach = np.zeros(1), np.zeros(1)
    flag = np.zeros(st.shape, dtype=np.uint8)
    L.ora_opcounts_reset()
    rc = L.ora_rk2_step(nx, ny, nz, st.dims, dx, dy, dz, gamma,
                        st.prim.ctypes.data_as(dp), st.bx.ctypes.data_as(dp),
                        st.by.ctypes.data_as(dp), st.bz.ctypes.data_as(dp), dt,
                        inv.ctypes.data_as(dp), mach.ctypes.data_as(dp),
                        flag.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), None, None, None)
    if rc < 0:
        raise ValueError("oracle rejected the box")
    out = (ctypes.c_ulonglong * 7)()
    L.ora_opcounts_get(out)
    return {k: int(out[i]) for i, k in enumerate(KINDS)}


def per_zone_update(m=32, seed=2024):
    """Operations per zone-update of the 3D random-field workload (HLLD,
    UCT_HLL, FP64), from boxes of m^3 and (2m)^3."""
    from pluto_amd import problems as P
    counts = {}
    for n in (m, 2 * m):
        st = P.random_mhd_3d(n, seed=seed)
        counts[n] = counted_step(st, 1e-4)
    a = {k: 2.0 * counts[2 * m][k] / (2 * m) ** 3 - counts[m][k] / m ** 3 for k in KINDS}
    flop = a["add"] + a["mul"] + a["div"] + a["sqrt"]
    return {
        "flop_per_zone_update": flop,
        "per_zone_update": a,
        "raw_counts": {str(n): counts[n] for n in counts},
        "flop_definition": "IEEE add/subtract, multiply, divide and square root of the "
                           "oracle's FP64 step, one each (no FMA: the path is compiled "
                           "without contraction); negation, fabs and comparisons are "
                           "counted apart and are not FLOPs",
        "workload": "3D random-field MHD state (seed %d), one RK2 step, HLLD + PLM(MC) + "
                    "CT(UCT_HLL), periodic; ops(n) = A n^3 + B n^2 fitted from n = %d, %d"
                    % (seed, m, 2 * m),
        "command": "python -m oracle.opcount --out profiles/flop_count_step.json",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--m", type=int, default=32)
    args = ap.parse_args()
    res = per_zone_update(args.m)
    text = json.dumps(res, indent=1, sort_keys=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(text)


if __name__ == "__main__":
    main()
