"""Branch signature and branch tape of the step (pgpu_set_branch_signature,
pgpu_set_branch_tape) and the 128^3 gradient problem checked with them.

The step is piecewise smooth: MC limiter cases, min/max of wave speeds,
HLLD regions and its degeneracy switch, UCT_HLL face-speed maxima and
MC_LIM2 slopes, ConsToPrim floors.  The discrete adjoint differentiates the
pieces the forward took; a central difference (J(x + h d) - J(x - h d)) / 2h
equals that derivative to O(h^2) only if x - h d, x and x + h d lie on one
piece.  The signature hashes every such comparison outcome of every cell,
stage and step, so equal signatures of the three runs say they did.  The
tape stores the outcomes of a run at x and imposes them on runs from other
states: the replayed objective is the smooth piece through x itself, whose
central differences are what the adjoint computes, with no switch inside
+-h.

The gradient problem: smooth 128^3 MHD state (rho = 1, p = 1,
B = (1, 0, 0) plus |k| <= 4 perturbations of rms 0.05 on rho, v and the
vector potential, seed 5), J = sum (rho_64 - rho_target)^2 after 64
fixed-dt RK2 steps, dJ/d(rho0, v0, A0) by the HIP adjoint with sqrt(N)
checkpointing, central differences at h = 1e-6 along 32 random directions
over every cell and control.  The difference
J(x + h d) - J(x - h d) is summed cell by cell as
(rho+ - rho-) (rho+ + rho- - 2 rho_target), which carries no cancellation
of J's ~1e4 magnitude.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def test_signature_leaves_the_step_alone_and_is_deterministic(native):
    from pluto_amd import problems as P
    _say("case: branch signature, random 32^3, 4 steps")
    st = P.random_mhd_3d(32, seed=11)

    def run(sig, state):
        dom = native.Domain(state.shape, 3, state.spacing(), 5.0 / 3.0)
        dom.set_state(state)
        dom.set_dt(2e-3, fixed=True)
        if sig:
            dom.set_branch_signature(True)
        dom.run(4)
        out = dom.get_state(state)
        s = dom.branch_signature() if sig else None
        dom.close()
        return out, s

    plain, _ = run(False, st)
    a, sa = run(True, st)
    b, sb = run(True, st)
    for x, y in ((plain.prim, a.prim), (plain.bx, a.bx), (plain.by, a.by), (plain.bz, a.bz)):
        assert np.array_equal(x, y)
    assert sa == sb and sa != 0
    # a perturbation of one cell's density large enough to move its
    # neighbours' limiter cases changes the signature
    moved = st.copy()
    moved.prim[0, 16, 16, 16] *= 1.5
    _, sm = run(True, moved)
    assert sm != sa
    # one ulp of one cell at a point far from every switch does not
    tiny = st.copy()
    tiny.prim[0, 3, 5, 7] = np.nextafter(tiny.prim[0, 3, 5, 7], 2.0)
    out_t, st_sig = run(True, tiny)
    assert not np.array_equal(out_t.prim, a.prim)
    _say("  signature %016x, moved %016x, ulp %016x" % (sa, sm, st_sig))
    assert st_sig == sa


def test_signature_refuses_what_it_does_not_cover(native):
    from pluto_amd import problems as P
    st = P.random_mhd_3d(16, seed=1)
    dom = native.Domain(st.shape, 3, st.spacing(), 5.0 / 3.0)
    dom.set_state(st)
    with pytest.raises(native.PGPUError):       # nothing recorded yet
        dom.branch_signature()
    dom.set_solver(native.SOLVER_HLLC)
    with pytest.raises(native.PGPUError):
        dom.set_branch_signature(True)
    dom.set_solver(native.SOLVER_HLLD)
    dom.set_branch_signature(True)
    dom.set_branch_signature(False)
    dom.close()
    dom = native.Domain(st.shape, 3, st.spacing(), 5.0 / 3.0,
                        precision=native.PGPU_PREC_FP32_FLUX)
    with pytest.raises(native.PGPUError):
        dom.set_branch_signature(True)
    dom.close()
    dom = native.Domain(st.shape, 3, st.spacing(), 5.0 / 3.0, bounds=["outflow"] * 6)
    with pytest.raises(native.PGPUError):
        dom.set_branch_signature(True)
    dom.close()


def test_tape_record_and_replay_reproduce_the_step(native):
    """Recording runs the modular kernels: bit-identical to the fused step.
    Replaying at the recorded state repeats it bit for bit; replaying at a
    state whose plain run keeps the signature is that plain run bit for bit;
    replaying at a state that moves switches differs from its plain run."""
    from pluto_amd import problems as P
    _say("case: branch tape, random 32^3, 4 steps")
    st = P.random_mhd_3d(32, seed=11)
    dom = native.Domain(st.shape, 3, st.spacing(), 5.0 / 3.0)
    dom.set_dt(2e-3, fixed=True)

    def plain(state, sig=False):
        dom.set_branch_tape(native.TAPE_OFF)
        dom.set_state(state)
        if sig:
            dom.set_branch_signature(True)
        dom.run(4)
        out = dom.get_state(state)
        s = dom.branch_signature() if sig else None
        dom.set_branch_signature(False)
        return out, s

    def same(a, b):
        return all(np.array_equal(x, y) for x, y in
                   ((a.prim, b.prim), (a.bx, b.bx), (a.by, b.by), (a.bz, b.bz)))

    ref, sig0 = plain(st, sig=True)
    dom.set_state(st)
    dom.set_branch_tape(native.TAPE_RECORD, 4)
    dom.run(4)
    assert same(dom.get_state(st), ref)

    def replay(state):
        dom.set_state(state)
        dom.set_branch_tape(native.TAPE_REPLAY)
        dom.run(4)
        return dom.get_state(state)

    assert same(replay(st), ref)
    tiny = st.copy()
    tiny.prim[0, 3, 5, 7] = np.nextafter(tiny.prim[0, 3, 5, 7], 2.0)
    pt, sig_t = plain(tiny, sig=True)
    dom.set_state(st)
    dom.set_branch_tape(native.TAPE_RECORD, 4)    # the tape again (plain() freed it)
    dom.run(4)
    assert sig_t == sig0 and not same(pt, ref)
    assert same(replay(tiny), pt)
    moved = st.copy()
    moved.prim[0, 16, 16, 16] *= 1.5
    pm, sig_m = plain(moved, sig=True)
    dom.set_state(st)
    dom.set_branch_tape(native.TAPE_RECORD, 4)
    dom.run(4)
    assert sig_m != sig0
    assert not same(replay(moved), pm)
    # past the recorded steps
    dom.set_state(st)
    dom.set_branch_tape(native.TAPE_REPLAY)
    with pytest.raises(native.PGPUError):
        dom.run(5)
    dom.set_branch_tape(native.TAPE_OFF)
    with pytest.raises(native.PGPUError):
        dom.set_branch_tape(native.TAPE_REPLAY)
    assert dom.scalars()["fail_cells"] == 0
    dom.close()


def test_gradient_128cube_vs_fd(native):
    """The 128^3, 64-step gradient against central differences at h = 1e-6
    along 32 seeded random directions over every cell of rho, v and A.

    Plain differences: each direction's +-h runs are screened with the
    branch signature (reported: at this size every such direction moves
    limiter or HLLD switches inside +-h, and the quotient then mixes one-sided
    slopes).  Differences of the branch-taped objective (the outcomes of the
    run at x imposed on the runs at x +- h d, x +- 2h d) are derivatives of
    the piece through x: the fourth-order central difference
    (8 (J(h) - J(-h)) - (J(2h) - J(-2h))) / 12h agrees with the adjoint to
    1e-6 on all 32 directions.  The two-point quotient is reported: its error
    is the h^2 truncation term (a random vector potential per edge moves B
    by ~h/dx, so J's third derivative along d is large), ~2.6e-6 at
    h = 1e-6, falling a hundredfold per tenfold smaller h.  Single control
    components are reported too (round-off limits their differences to
    ~1e-6 at h = 1e-6)."""
    from pluto_amd import problems as P
    from pluto_amd.adjoint import Stepper, cfl_dt, controls_to_state, rollout
    torch = _torch()
    n, nsteps, h = 128, 64, 1e-6
    _say("case: gradient problem 128^3, 64 steps, FD at h=1e-6, signatures and branch tape")
    rho, v, a = P.smooth_mhd_controls(n, seed=5)
    st = P.smooth_mhd_3d(n, seed=5)
    dt = cfl_dt(st, 0.4)
    d = st.spacing()
    stepper = Stepper(st.shape, d, dt)
    dom = stepper.dom
    dev = dict(dtype=torch.float64, device="cuda")
    p_t = torch.ones((n, n, n), **dev)
    rho_t = torch.tensor(P.smooth_mhd_3d(n, seed=6).prim[0], **dev)
    ctl = [torch.tensor(rho, **dev), torch.tensor(v, **dev), torch.tensor(np.stack(a), **dev)]

    leaves = [c.clone().requires_grad_(True) for c in ctl]
    prim, faces = controls_to_state(*leaves, p_t, (1.0, 0.0, 0.0), d)
    op, _ = rollout(prim.contiguous(), faces.contiguous(), stepper, nsteps)
    J = torch.sum((op[0] - rho_t) ** 2)
    J.backward()
    g = [x.grad.detach() for x in leaves]
    rho_plain = op[0].detach().clone()

    def final_rho(r, vv, aa, mode):
        """mode "sig": plain run, returns its signature; "replay": taped."""
        with torch.no_grad():
            pr, fa = controls_to_state(r, vv, aa, p_t, (1.0, 0.0, 0.0), d)
            stepper.load(pr.contiguous(), fa.contiguous())
            if mode == "sig":
                dom.set_branch_signature(True)
            else:
                dom.set_branch_tape(native.TAPE_REPLAY)
            stepper.run(nsteps)
            out, _ = stepper.save()
            sig = dom.branch_signature() if mode == "sig" else None
            dom.set_branch_signature(False)
            return out[0].clone(), sig

    def fd_of(rp, rm):
        return float(torch.sum((rp - rm) * (rp + rm - 2.0 * rho_t))) / (2 * h)

    shapes = [c.shape for c in ctl]

    def direction(t):
        rng = np.random.default_rng([2025, t])
        return [torch.tensor(rng.standard_normal(s), **dev) for s in shapes]

    # plain differences of the first 8 directions, screened by signature
    _, sig0 = final_rho(*ctl, "sig")
    plain_err, unbroken = [], 0
    for t in range(8):
        dirs = direction(t)
        ad = sum(float(torch.sum(gi * di)) for gi, di in zip(g, dirs))
        rp, sp = final_rho(*[c + h * e for c, e in zip(ctl, dirs)], "sig")
        rm, sm = final_rho(*[c - h * e for c, e in zip(ctl, dirs)], "sig")
        ok = sp == sig0 and sm == sig0
        unbroken += ok
        plain_err.append(abs(fd_of(rp, rm) - ad) / abs(ad))
        _say("  dir %2d: adjoint %+.12e plain fd rel %.2e, signature kept %s"
             % (t, ad, plain_err[-1], ok))

    # the tape of the run at x, then differences of the taped objective
    with torch.no_grad():
        pr, fa = controls_to_state(*ctl, p_t, (1.0, 0.0, 0.0), d)
        stepper.load(pr.contiguous(), fa.contiguous())
        dom.set_branch_tape(native.TAPE_RECORD, nsteps)
        stepper.run(nsteps)
        out, _ = stepper.save()
    assert torch.equal(out[0], rho_plain)          # taped kernels == fused kernels
    taped_err, two_point = [], []
    for t in range(32):
        dirs = direction(t)
        ad = sum(float(torch.sum(gi * di)) for gi, di in zip(g, dirs))
        q = {}
        for m in (1, 2):
            q[m] = fd_of(final_rho(*[c + m * h * e for c, e in zip(ctl, dirs)], "replay")[0],
                         final_rho(*[c - m * h * e for c, e in zip(ctl, dirs)], "replay")[0]) / m
        taped = (4.0 * q[1] - q[2]) / 3.0          # fourth-order central difference
        taped_err.append(abs(taped - ad) / abs(ad))
        two_point.append(abs(q[1] - ad) / abs(ad))
        _say("  dir %2d: adjoint %+.12e taped fd %+.12e rel %.2e (two-point %.2e)"
             % (t, ad, taped, taped_err[-1], two_point[-1]))
    comp_err = []
    rng = np.random.default_rng(7)
    for t in range(9):
        which = t % 3
        idx = tuple(int(i) for i in rng.integers(0, n, 3))
        dirs = [torch.zeros_like(c) for c in ctl]
        dirs[which][idx if which == 0 else (int(rng.integers(0, 3)),) + idx] = 1.0
        ad = sum(float(torch.sum(gi * di)) for gi, di in zip(g, dirs))
        taped = fd_of(final_rho(*[c + h * e for c, e in zip(ctl, dirs)], "replay")[0],
                      final_rho(*[c - h * e for c, e in zip(ctl, dirs)], "replay")[0])
        comp_err.append(abs(taped - ad) / abs(ad))
        _say("  component %d of control %d: adjoint %+.12e taped fd %+.12e rel %.2e"
             % (t, which, ad, taped, comp_err[-1]))
    assert dom.scalars()["fail_cells"] == 0        # no ConsToPrim floor (not taped)
    dom.set_branch_tape(native.TAPE_OFF)
    stepper.close()
    _say("  32 directions: taped fd max rel err %.2e (median %.2e; two-point max %.2e); plain "
         "fd on the first 8: signature kept on %d, max rel err %.2e; single components: taped "
         "max rel err %.2e" % (max(taped_err), float(np.median(taped_err)), max(two_point),
                               unbroken, max(plain_err), max(comp_err)))
    assert max(taped_err) <= 1e-6
