"""The operation-counting build of the oracle (oracle/libopcount.so) behind
the algorithmic FLOP count of a zone-update: it computes exactly what the
oracle computes, and its committed count (profiles/flop_count_step.json) is
what it counts now."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT


def test_counted_oracle_computes_the_oracle_step(oracle_lib):
    from oracle import opcount
    from pluto_amd import problems as P
    st = P.random_mhd_3d(12, seed=5)
    a, b = st.copy(), st.copy()
    oracle_lib.rk2_step(a, 2e-3)
    c = opcount.counted_step(b, 2e-3)
    for x, y in ((a.prim, b.prim), (a.bx, b.bx), (a.by, b.by), (a.bz, b.bz)):
        assert np.array_equal(x, y)
    assert c["add"] > 0 and c["mul"] > 0 and c["div"] > 0 and c["sqrt"] > 0


def test_branch_free_counts_follow_the_row_ranges():
    """Square roots do not depend on branches: 48 per zone-update (2 stages x
    3 sweeps x 8 per interface) plus PLUTO's one extra interface per row."""
    from oracle import opcount
    from pluto_amd import problems as P
    for n in (8, 12):
        c = opcount.counted_step(P.random_mhd_3d(n, seed=1), 1e-3)
        assert c["sqrt"] == 48 * n ** 3 + 48 * n ** 2


def test_committed_flop_count_is_current():
    from oracle import opcount
    path = os.path.join(ROOT, "profiles", "flop_count_step.json")
    with open(path) as f:
        ref = json.load(f)
    now = opcount.per_zone_update()
    assert now["flop_per_zone_update"] == pytest.approx(ref["flop_per_zone_update"], rel=0, abs=0)
    assert now["per_zone_update"] == ref["per_zone_update"]
