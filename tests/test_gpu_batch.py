"""Batched patches: every patch of a batch evolves bit for bit like a single
domain of its size (which is itself bit-identical to the oracle), whatever
the launch shape, with its own CFL time step."""
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


EDGES = [16, 12, 8, 5, 3, 7, 16, 9, 4, 11]


def _states():
    from pluto_amd import problems as P
    return [P.random_mhd_3d(n, seed=100 + p) for p, n in enumerate(EDGES)]


def _single(native, st, nsteps):
    dom = native.Domain.from_state(st)
    dom.set_dt(1e-4, 0.4, 1.1)
    dom.run(nsteps)
    out = dom.get_state(st)
    sc = dom.scalars()
    dom.close()
    return out, sc


def test_batch_matches_single_domains(native):
    _say("case: batch of %d patches %s, 6 CFL steps, vs single domains" % (len(EDGES), EDGES))
    states = _states()
    b = native.Batch(EDGES)
    for p, st in enumerate(states):
        b.set_patch(p, st)
    b.set_dt(1e-4, 0.4, 1.1)
    b.run(6)
    b.sync()
    for p, st in enumerate(states):
        ref, rsc = _single(native, st, 6)
        out = b.get_patch(p)
        assert_states_equal(out, ref, "patch %d (edge %d)" % (p, EDGES[p]))
        sc = b.scalars(p)
        assert sc["step"] == 6
        assert sc["time"] == rsc["time"], (p, sc["time"], rsc["time"])
        assert sc["dt"] == rsc["dt"]
        assert sc["maxmach"] == rsc["maxmach"]
    b.close()


@pytest.mark.parametrize("threads,groups", [(64, 1), (128, 3), (512, 1000)])
def test_batch_launch_shape_does_not_change_results(native, threads, groups):
    _say("case: batch launch shape %d threads x %d groups" % (threads, groups))
    states = _states()
    ref = native.Batch(EDGES)
    alt = native.Batch(EDGES)
    for p, st in enumerate(states):
        ref.set_patch(p, st)
        alt.set_patch(p, st)
    for b in (ref, alt):
        b.set_dt(1e-4, 0.4, 1.1)
    alt.set_launch(threads, groups)
    ref.run(3)
    alt.run(3)
    for p in range(len(EDGES)):
        assert_states_equal(alt.get_patch(p), ref.get_patch(p), "patch %d" % p)
    ref.close()
    alt.close()


def test_batch_tune_and_graph_replay(native):
    _say("case: batch autotuner, then graph replays")
    states = _states()
    b = native.Batch(EDGES)
    for p, st in enumerate(states):
        b.set_patch(p, st)
    b.set_dt(1e-4, 0.4, 1.1)
    t, g, ms = b.tune(1)
    assert t in (128, 256, 512, 1024) and g >= 1 and ms > 0
    assert b.launch() == (t, g)
    b.run(2)
    steps = b.scalars(0)["step"]
    assert steps >= 3
    for p in (0, 4, 9):
        ref, _ = _single(native, states[p], steps)
        assert_states_equal(b.get_patch(p), ref, "tuned patch %d" % p)
    b.close()


def test_batch_rejects_bad_edges(native):
    with pytest.raises(native.PGPUError):
        native.Batch([8, 2])


@pytest.mark.parametrize("threads,tile", [(256, (16, 4, 4)), (256, (8, 8, 4)), (128, (4, 8, 4)),
                                          (512, (64, 8, 1)), (1024, (16, 8, 8))])
def test_batch_cell_tile_does_not_change_results(native, oracle_lib, threads, tile):
    """Tiled cell passes (boxes larger than some patches, partial boxes at
    patch edges) give the same bits as the linear order and the oracle."""
    _say("case: batch cell tile %s at %d threads" % (tile, threads))
    states = _states()
    b = native.Batch(EDGES)
    for p, st in enumerate(states):
        b.set_patch(p, st)
    b.set_dt(1e-4, 0.4, 1.1)
    b.set_launch(threads, 7)
    b.set_tile(*tile)
    assert b.tile() == tuple(tile)
    b.run(3)
    b.sync()
    for p, st in enumerate(states):
        ref = st.copy()
        oracle_lib.run(ref, 3, 1e-4, cfl=0.4, cfl_max_var=1.1)
        assert_states_equal(b.get_patch(p), ref, "patch %d (edge %d), tile %s" % (p, EDGES[p], tile))
    with pytest.raises(native.PGPUError):
        b.set_tile(3, 3, 3)                 # not the workgroup's threads
    b.set_tile()
    assert b.tile() == (0, 0, 0)
    b.close()


def test_batch_tuner_picks_launch_and_tile(native):
    _say("case: batch autotuner (launch shapes, then cell tiles)")
    states = _states()
    b = native.Batch(EDGES)
    for p, st in enumerate(states):
        b.set_patch(p, st)
    b.set_dt(1e-4, 0.4, 1.1)
    threads, groups, ms = b.tune(1)
    assert (threads, groups) == b.launch() and ms > 0
    tile = b.tile()
    assert tile == (0, 0, 0) or tile[0] * tile[1] * tile[2] == threads
    b.close()


STACK_EDGES = [8, 16, 8, 12, 16, 8, 4, 24, 6, 16]


def test_stacked_mode_matches_persistent_and_oracle(native, oracle_lib):
    """Patches of even edges stacked by size through the fused kernels (z
    periodic within each patch, per-patch dt): bit-identical to the
    persistent patch kernel and to the oracle run of each patch, over CFL
    steps where the patches' time steps differ."""
    from pluto_amd import problems as P
    _say("case: stacked batch %s vs persistent and oracle, 4 CFL steps" % STACK_EDGES)
    states = [P.random_mhd_3d(n, seed=300 + p) for p, n in enumerate(STACK_EDGES)]
    st_b = native.Batch(STACK_EDGES)
    pe_b = native.Batch(STACK_EDGES)
    pe_b.set_mode(0)
    assert st_b.mode() == (1, len(STACK_EDGES))
    assert pe_b.mode() == (0, 0)
    # cfl_max_var 10: dt grows until each patch's CFL limit binds
    for b in (st_b, pe_b):
        for p, st in enumerate(states):
            b.set_patch(p, st)
        b.set_dt(1e-4, 0.4, 10.0)
        b.run(4)
        b.sync()
    dts = set()
    for p, st in enumerate(states):
        ref = st.copy()
        oracle_lib.run(ref, 4, 1e-4, cfl=0.4, cfl_max_var=10.0)
        out = st_b.get_patch(p)
        assert_states_equal(out, ref, "stacked patch %d (edge %d)" % (p, STACK_EDGES[p]))
        assert_states_equal(pe_b.get_patch(p), out, "persistent patch %d" % p)
        s1, s0 = st_b.scalars(p), pe_b.scalars(p)
        assert s1 == s0, (p, s1, s0)
        assert s1["step"] == 4
        dts.add(s1["dt"])
    assert len(dts) > 1          # the patches keep their own time steps
    st_b.close()
    pe_b.close()


def test_stacked_mode_switches(native):
    """Mixed batches stack only the even edges; switching the x/y flux kernel
    off the row kernel makes the next capture persistent; bad modes refused."""
    b = native.Batch(EDGES)
    n_even = sum(1 for n in EDGES if n % 2 == 0 and n >= 4)
    assert b.mode() == (1, n_even)
    native.set_flux_kernel(0, 0)
    try:
        b.set_mode(1)
        assert b.mode() == (1, 0)
    finally:
        native.set_flux_kernel()
    b.set_mode(1)
    assert b.mode() == (1, n_even)
    with pytest.raises(native.PGPUError):
        b.set_mode(2)
    b.close()
