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


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_run(name):
    """A stored reference run: (initial MHDState, {step: (meta, MHDState)}, run kwargs)."""
    import json
    from pluto_amd.problems import MHDState
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    dims = int(z["dims"][0])
    length = tuple(float(x) for x in z["length"])
    init = MHDState(z["init_prim"].copy(), z["init_bx"].copy(), z["init_by"].copy(),
                    z["init_bz"].copy(), dims, length)
    steps = {}
    for s, h in meta["steps"].items():
        s = int(s)
        steps[s] = (h, MHDState(z["s%d_prim" % s], z["s%d_bx" % s], z["s%d_by" % s],
                                z["s%d_bz" % s], dims, length))
    return init, steps, meta["run"]


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
