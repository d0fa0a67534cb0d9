"""The oracle's Boundary (oracle/boundary.py, a numpy restatement of
boundary.c and ct_fill_mag_field.c) against the compiled reference's own
Boundary (oracle/_ref/libplutoref_rows.so) on the CPU, bit for bit, for
every side type and side mix on random ghosted boxes (every element random,
ghosts and -1 face planes included; random face areas).  A planted sign
error in the oracle's FlipSign is caught.
"""
import ctypes
import os

import numpy as np
import pytest

from pluto_amd import plutoabi as A

OUT, REF, AXI, EQT, PER = 1, 2, 3, 4, 5
CASES = {
    "outflow": (OUT,) * 6,
    "reflective": (REF,) * 6,
    "axisymmetric": (AXI,) * 6,
    "eqtsymmetric": (EQT,) * 6,
    "periodic": (PER,) * 6,
    "mixed": (OUT, REF, PER, PER, EQT, OUT),
    "mixed_internal": (REF, 0, OUT, AXI, 0, PER),
}


@pytest.fixture(scope="module")
def ref():
    from oracle import refrun
    path = os.path.join(refrun.REF_DIR, "libplutoref_rows.so")
    if not os.path.exists(path):
        pytest.skip("reference row library not built (make -C oracle/ref)")
    return A.bind_boundary(ctypes.CDLL(path))


def _box(n, ng, bounds, seed):
    from pluto_amd import problems as P
    st = P.random_mhd_3d(max(n), seed=seed)
    st.prim = np.ascontiguousarray(st.prim[:, :n[2], :n[1], :n[0]])
    st.bx, st.by, st.bz = (np.ascontiguousarray(b[:n[2], :n[1], :n[0]]) for b in (st.bx, st.by, st.bz))
    box = A.PlutoBox(st, ng=ng)
    rng = np.random.default_rng(seed)
    box.vc[:] = rng.normal(size=box.vc.shape)
    for a in range(3):
        box.vs[a][:] = rng.normal(size=box.vs[a].shape)
    area = []
    for a, lo in enumerate(((0, 0, -1), (0, -1, 0), (-1, 0, 0))):
        arr = box.keep(rng.uniform(0.5, 1.5, size=box.vs[a].shape))
        area.append(arr)
        box.grid.A[a] = A.table3(box.keep, arr, lo)
    for a in range(3):
        box.grid.lbound[a], box.grid.rbound[a] = bounds[2 * a], bounds[2 * a + 1]
    return box, area


def _run_ref(ref, box, idim):
    g = box.grid
    for a, (b, e, n, t) in enumerate(zip(("IBEG", "JBEG", "KBEG"), ("IEND", "JEND", "KEND"),
                                         ("NX1", "NX2", "NX3"), ("NX1_TOT", "NX2_TOT", "NX3_TOT"))):
        ctypes.c_long.in_dll(ref, b).value = g.lbeg[a]
        ctypes.c_long.in_dll(ref, e).value = g.lend[a]
        ctypes.c_long.in_dll(ref, n).value = g.np_int[a]
        ctypes.c_long.in_dll(ref, t).value = g.np_tot[a]
    ref.Boundary(ctypes.byref(box.data), idim, ctypes.byref(box.grid))


def _run_oracle(box, area, bounds, idim, flip=None):
    from oracle import boundary as B
    vc = [box.vc[q].copy() for q in range(8)]
    vs = [v.copy() for v in box.vs]
    g = box.grid
    beg = [g.lbeg[a] for a in range(3)]
    end = [g.lend[a] for a in range(3)]
    saved = B.flip_sign
    if flip is not None:
        B.flip_sign = flip
    try:
        B.boundary(vc, vs, area, bounds, beg, end, idim)
    finally:
        B.flip_sign = saved
    return vc + vs


@pytest.mark.parametrize("idim", [-1, 0, 1, 2], ids=["all", "x1", "x2", "x3"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_boundary_vs_reference(ref, case, idim):
    box, area = _box((9, 8, 7), 2, CASES[case], seed=5)
    mine = _run_oracle(box, area, CASES[case], idim)
    _run_ref(ref, box, idim)
    theirs = [box.vc[q] for q in range(8)] + list(box.vs)
    for q in range(11):
        assert np.array_equal(mine[q], theirs[q]), (case, idim, q)


def test_oracle_boundary_three_ghosts(ref):
    box, area = _box((7, 5, 4), 3, CASES["mixed"], seed=9)
    mine = _run_oracle(box, area, CASES["mixed"], -1)
    _run_ref(ref, box, -1)
    theirs = [box.vc[q] for q in range(8)] + list(box.vs)
    for q in range(11):
        assert np.array_equal(mine[q], theirs[q]), q


def test_planted_sign_error_is_caught(ref):
    from oracle import boundary as B
    good = B.flip_sign

    def bad(d, btype):
        s = good(d, btype)
        s[B.BX1 + (d + 1) % 3] *= -1.0     # a tangential field flipped by mistake
        return s

    box, area = _box((9, 8, 7), 2, CASES["reflective"], seed=5)
    mine = _run_oracle(box, area, CASES["reflective"], -1, flip=bad)
    _run_ref(ref, box, -1)
    theirs = [box.vc[q] for q in range(8)] + list(box.vs)
    assert not all(np.array_equal(mine[q], theirs[q]) for q in range(11))
