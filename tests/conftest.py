import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def state_diff(a, b):
    """Count of differing elements per array (bitwise, -0.0 == 0.0)."""
    out = {}
    for name, x, y in (("prim", a.prim, b.prim), ("bx", a.bx, b.bx),
                       ("by", a.by, b.by), ("bz", a.bz, b.bz)):
        out[name] = int(np.sum(x != y))
    return out


def assert_states_equal(a, b, what=""):
    d = state_diff(a, b)
    assert all(v == 0 for v in d.values()), "%s differs: %s" % (what, d)


def max_rel_diff(a, b):
    """Max over arrays of |a-b| / max(|b|) per field."""
    worst = 0.0
    for x, y in ((a.prim, b.prim), (a.bx, b.bx), (a.by, b.by), (a.bz, b.bz)):
        if x.ndim == 4:
            for q in range(x.shape[0]):
                s = np.max(np.abs(y[q]))
                if s > 0:
                    worst = max(worst, float(np.max(np.abs(x[q] - y[q])) / s))
        else:
            s = np.max(np.abs(y))
            if s > 0:
                worst = max(worst, float(np.max(np.abs(x - y)) / s))
    return worst


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    O.lib()
    return O
