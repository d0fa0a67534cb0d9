"""The benchmark workloads at their full sizes.

* 3D 256^3 random-field MHD (seed 2024; the headline workload): 3 CFL steps
  bit for bit against the oracle (OpenMP), then to step 100 on the device
  with div.B at round-off and sum(rho), sum(m), sum(E) conserved to
  round-off (periodic box).
* 4096 independent patches, edges uniform in {16, 24, 32, 48, 64} (seed
  2024), each a random-field state at its own size: 2 steps of the
  persistent batch, and 64 sampled patches (every edge size) bit for bit
  against the oracle run on that patch alone (states and dt).
* 512^3 random-field state, 50 steps of the FP32-flux path against the FP64
  path (each with its own CFL control): per-field relative L1 <= 1e-6 (the
  precision split's error bar), L-infinity reported.
"""
import math

import numpy as np
import pytest

from conftest import assert_states_equal

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


def _totals(st):
    """sum rho, sum m (3), sum E, and sum |m| as the momentum scale."""
    rho, v, b, p = st.prim[0], st.prim[1:4], st.prim[4:7], st.prim[7]
    m = rho[None] * v
    E = p / (5.0 / 3.0 - 1.0) + 0.5 * np.sum(rho[None] * v * v, axis=0) + 0.5 * np.sum(b * b, axis=0)
    return (math.fsum(rho.ravel()), [math.fsum(x.ravel()) for x in m], math.fsum(E.ravel()),
            float(np.abs(m).sum()))


@pytest.mark.timeout(600)
def test_headline_256cube_bitexact_then_100_steps(native, oracle_lib):
    from pluto_amd import problems as P
    n = 256
    _say("case: 256^3 random field (seed 2024): 3 CFL steps vs the oracle, then to step 100")
    st = P.random_mhd_3d(n, seed=2024)
    t0 = _totals(st)
    ref = st.copy()
    dts, mach = oracle_lib.run(ref, 3, 1e-4, cfl=0.4, cfl_max_var=1.1)
    dom = native.Domain.from_state(st)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(3)
    assert_states_equal(dom.get_state(st), ref, "256^3 after 3 steps")
    hdt, _, hmach = dom.history(3)
    assert np.array_equal(hdt, dts) and np.array_equal(hmach, mach)
    del ref
    dom.run(97)
    dom.sync()
    out = dom.get_state(st)
    assert dom.scalars()["step"] == 100
    bmax = max(np.abs(b).max() for b in (out.bx, out.by, out.bz))
    divb = np.abs(P.divb(out)).max() / n / bmax
    _say("  div.B/|B|max after 100 steps: %.3e" % divb)
    assert divb <= 1e-13
    t1 = _totals(out)
    assert abs(t1[0] - t0[0]) <= 1e-13 * abs(t0[0])
    for c in range(3):
        assert abs(t1[1][c] - t0[1][c]) <= 1e-13 * t0[3], c
    assert abs(t1[2] - t0[2]) <= 1e-13 * abs(t0[2])
    assert not np.array_equal(out.prim, st.prim)
    dom.close()


def _batch_patches(npatch=4096, sizes=(16, 24, 32, 48, 64), seeds_per_size=8, seed=2024):
    """The patch workload: edges uniform in sizes (seeded), each patch a
    seeded periodic translate of one of seeds_per_size random-field states
    made at its own size.  Returns the edges and make(p), which builds
    patch p's state (the 4096 states are not held at once)."""
    from pluto_amd import problems as P
    rng = np.random.default_rng(seed)
    edges = [int(x) for x in rng.choice(sizes, npatch)]
    base = {n: [P.random_mhd_3d(n, seed=seed + 1000 * n + s) for s in range(seeds_per_size)]
            for n in sizes}
    recipe = []
    for n in edges:
        which = int(rng.integers(seeds_per_size))
        sh = tuple(int(x) for x in rng.integers(0, n, 3))
        recipe.append((which, sh))

    def make(p):
        n = edges[p]
        which, sh = recipe[p]
        st = base[n][which]
        prim = np.roll(st.prim, sh, axis=(1, 2, 3))
        faces = [np.roll(a, sh, axis=(0, 1, 2)) for a in (st.bx, st.by, st.bz)]
        return P.MHDState(prim, *faces, 3, (1.0, 1.0, 1.0))

    return edges, make


@pytest.mark.timeout(600)
def test_batch_4096_patches_vs_oracle(native, oracle_lib):
    _say("case: 4096-patch batch (edges 16..64, seed 2024), 2 steps, 64 patches vs the oracle")
    edges, make = _batch_patches()
    b = native.Batch(edges)
    for p in range(len(edges)):
        b.set_patch(p, make(p))
    b.set_dt(1e-4, 0.4, 1.1)
    b.tune(1)                     # the autotuner's trial steps advance the patches:
    for p in range(len(edges)):
        b.set_patch(p, make(p))   # start again from the workload's state
    b.set_dt(1e-4, 0.4, 1.1)
    b.run(2)
    b.sync()
    # 64 patches, the same number of each edge size (seeded)
    rng = np.random.default_rng(7)
    sizes = sorted(set(edges))
    picks = []
    for q, n in enumerate(sizes):
        idx = [p for p, e in enumerate(edges) if e == n]
        k = 64 // len(sizes) + (1 if q < 64 % len(sizes) else 0)
        picks += [int(x) for x in rng.choice(idx, k, replace=False)]
    assert len(picks) == 64 and {edges[p] for p in picks} == set(sizes)
    for p in sorted(picks):
        ref = make(p)
        dts, mach = oracle_lib.run(ref, 2, 1e-4, cfl=0.4, cfl_max_var=1.1)
        assert_states_equal(b.get_patch(p), ref, "patch %d (edge %d)" % (p, edges[p]))
        sc = b.scalars(p)
        assert sc["step"] == 2 and sc["time"] == dts[0] + dts[1], p
        assert sc["maxmach"] == mach[1], p
    b.close()


@pytest.mark.timeout(900)
def test_fp32_flux_512cube_50_steps(native):
    from pluto_amd import problems as P
    n = 512
    _say("case: 512^3 random field, 50 steps, FP32 flux vs FP64 (own CFL control each)")
    st = P.random_mhd_3d(n, seed=2024)
    d64 = native.Domain.from_state(st)
    d32 = native.Domain.from_state(st, precision=native.PGPU_PREC_FP32_FLUX)
    del st
    for d in (d64, d32):
        d.set_dt(1e-4, 0.4, 1.1)
        d.run(50)
        d.sync()
    err = d32.error_vs(d64)
    for f, (l1, linf) in err.items():
        _say("  %-3s L1 %.3e  Linf %.3e" % (f, l1, linf))
        assert l1 <= 1e-6, (f, l1)
    s64, s32 = d64.scalars(), d32.scalars()
    assert s64["step"] == s32["step"] == 50
    assert abs(s32["time"] / s64["time"] - 1.0) < 1e-5
    d64.close()
    d32.close()
