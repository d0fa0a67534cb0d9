"""CPU restatement of PLUTO's Boundary() for Cartesian MHD (numpy).

TEST INFRASTRUCTURE ONLY: the oracle the device ghost fill is checked
against.  Imported by tests/ -- never by the product package.

Arrays are PLUTO's, as numpy blocks in [k][j][i] order: eight cell-centred
primitives of shape (t2, t1, t0), and three staggered components whose own
axis holds the face plane -1 first (shape t+1 along it), with face areas of
the same shapes.  beg/end are IBEG..IEND per axis (x, y, z).

  boundary()            boundary.c:40-315 (the side loop, 5a-5b)
  _centre/_face boxes   boundary.c:144-165 (RBoxDefine of the four boxes)
  _fill                 OutflowBound, PeriodicBound, ReflectiveBound
                        (boundary.c:440-564), marched in BOX_LOOP order
                        (macros.h:91-94)
  flip_sign             FlipSign (boundary.c:318-430), MHD, 3 components
  _fill_mag             FillMagneticField (MHD/CT/ct_fill_mag_field.c:38-140)
  _average_normal       CT_AverageNormalMagField, Cartesian
                        (MHD/CT/ct_field_average.c:134-270)
"""

from __future__ import annotations

import numpy as np

OUTFLOW, REFLECTIVE, AXISYMMETRIC, EQTSYMMETRIC, PERIODIC = 1, 2, 3, 4, 5
VX1, BX1 = 1, 4


def flip_sign(d, btype):
    s = [1.0] * 8
    n, p, t = d, (1 if d == 0 else 0), (1 if d == 2 else 2)
    s[VX1 + n] = -1.0
    s[BX1 + n] = -1.0
    if btype == AXISYMMETRIC:
        s[VX1 + t] = -1.0
        s[BX1 + t] = -1.0
    if btype == EQTSYMMETRIC:
        s[BX1 + n] = 1.0
        s[BX1 + p] = -1.0
        s[BX1 + t] = -1.0
    return s


class _View:
    """PLUTO index (i, j, k) -> numpy element, with the -1 plane of `ext`."""

    def __init__(self, arr, ext):
        self.a, self.o = arr, [1 if ext == a else 0 for a in range(3)]

    def __getitem__(self, p):
        return self.a[p[2] + self.o[2], p[1] + self.o[1], p[0] + self.o[0]]

    def __setitem__(self, p, v):
        self.a[p[2] + self.o[2], p[1] + self.o[1], p[0] + self.o[0]] = v


def _range(lo, hi):
    step = 1 if lo <= hi else -1
    return range(lo, hi + step, step)


def _fill(view, box, side, kind, sgn, beg, end, n):
    d, hi_side = side // 2, side & 1
    for k in _range(box[2][0], box[2][1]):
        for j in _range(box[1][0], box[1][1]):
            for i in _range(box[0][0], box[0][1]):
                p = [i, j, k]
                q = list(p)
                m = p[d]
                if kind == OUTFLOW:
                    q[d] = end[d] if hi_side else beg[d]
                elif kind == PERIODIC:
                    q[d] = m - n[d] if hi_side else m + n[d]
                else:
                    q[d] = 2 * end[d] - m + 1 if hi_side else 2 * beg[d] - m - 1
                v = view[q]
                view[p] = sgn * v if kind == REFLECTIVE else v


def _fill_mag(bs, area, side, t, beg, end):
    d, hi_side = side // 2, side & 1
    others = [a for a in range(3) if a != d]
    ranges = [range(0, t[a]) for a in range(3)]
    ranges[d] = range(end[d] + 1, t[d]) if hi_side else range(beg[d] - 1, -1, -1)
    for k in ranges[2]:
        for j in ranges[1]:
            for i in ranges[0]:
                p = [i, j, k]
                dB = []
                for e in others:
                    m = list(p)
                    m[e] -= 1
                    dB.append(area[e][p] * bs[e][p] - area[e][m] * bs[e][m])
                m = list(p)
                m[d] -= 1
                if hi_side:
                    bs[d][p] = (area[d][m] * bs[d][m] - (dB[0] + dB[1])) / area[d][p]
                else:
                    bs[d][m] = (area[d][p] * bs[d][p] + dB[0] + dB[1]) / area[d][m]


def _average_normal(vc, bs, side, t, beg, end):
    d, hi_side = side // 2, side & 1
    ranges = [range(0, t[a]) for a in range(3)]
    ranges[d] = range(end[d] + 1, t[d]) if hi_side else range(0, beg[d])
    bc = _View(vc[BX1 + d], -1)
    for k in ranges[2]:
        for j in ranges[1]:
            for i in ranges[0]:
                p = [i, j, k]
                m = list(p)
                m[d] -= 1
                bc[p] = 0.5 * (bs[d][p] + bs[d][m])


def boundary(vc, vs, area, types, beg, end, idim=-1):
    """Fill the ghost zones of vc (list of 8 arrays) and vs (3 arrays) in
    place for the sides of idim (0, 1, 2 or -1 = all) with types[6] (X1_BEG,
    X1_END, ... X3_END; 0 = internal side, skipped)."""
    t = [vc[0].shape[2], vc[0].shape[1], vc[0].shape[0]]
    n = [end[a] - beg[a] + 1 for a in range(3)]
    bs = [_View(vs[a], a) for a in range(3)]
    ar = [_View(area[a], a) for a in range(3)]
    sides = range(6) if idim < 0 else (2 * idim, 2 * idim + 1)
    for side in sides:
        btype = types[side]
        if btype == 0:
            continue
        d, hi_side = side // 2, side & 1
        kind = OUTFLOW if btype == OUTFLOW else PERIODIC if btype == PERIODIC else REFLECTIVE
        sgn = flip_sign(d, btype)
        centre = [[0, t[a] - 1] for a in range(3)]
        centre[d] = [end[d] + 1, t[d] - 1] if hi_side else [beg[d] - 1, 0]
        for q in range(8):
            _fill(_View(vc[q], -1), centre, side, kind, sgn[q], beg, end, n)
        for a in range(3):
            if a == d and kind != PERIODIC:
                continue
            box = [list(r) for r in centre]
            if box[a][0] < box[a][1]:
                box[a][0] -= 1
            else:
                box[a][1] -= 1
            _fill(bs[a], box, side, kind, sgn[BX1 + a], beg, end, n)
        if kind != PERIODIC:
            _fill_mag(bs, ar, side, t, beg, end)
        if kind == OUTFLOW:
            _average_normal(vc, bs, side, t, beg, end)
