"""Discrete adjoint of the RK2 step (HIP) against the forward model.

Point level: HLLD's hand-written adjoint is the exact transpose of its
tangent model (forward-mode dual numbers through the same template the
forward kernels use): <lam, J t> = <J^T lam, t> to round-off on the
reference's own HLLD rows, which cover all six regions and the degeneracy
branch.  The tangent's values are the oracle's HLLD bit for bit and its
derivatives match central differences of the oracle.

Step level: <dJ/dx, d> from pgpu_adjoint_step against central differences
(J(x + h d) - J(x - h d)) / 2h of the forward path, which is bit-identical
to the oracle, for linear functionals of the final state after one and
several steps and for the config-5 style objective sum (rho - rho_t)^2 with
respect to rho, v and the vector potential; checkpointed and
uncheckpointed rollouts agree to round-off.
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

OUT_FLUX = [0, 1, 2, 3, 7]     # rho, mx, my, mz, E columns of the reference flux rows


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def native():
    from pluto_amd import native as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


def _p(a):
    import ctypes
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _tangent(N, d, vl, vr, tl, tr, gamma=5.0 / 3.0):
    n = vl.shape[0]
    out, dout = np.zeros((n, 8)), np.zeros((n, 8))
    N.check(N.lib().pgpu_hlld_tangent_rows(d, n, gamma, _p(vl), _p(vr), _p(tl), _p(tr),
                                            _p(out), _p(dout)))
    return out, dout


def _adjoint(N, d, vl, vr, lam, gamma=5.0 / 3.0):
    n = vl.shape[0]
    al, ar = np.zeros((n, 8)), np.zeros((n, 8))
    N.check(N.lib().pgpu_hlld_adjoint_rows(d, n, gamma, _p(vl), _p(vr), _p(lam), _p(al), _p(ar)))
    return al, ar


def _oracle_out(oracle_lib, d, vl, vr):
    flux, press, sl, sr, _, _ = oracle_lib.hlld_rows(d, vl, vr)
    return np.concatenate([flux[:, OUT_FLUX], press[:, None], sl[:, None], sr[:, None]], axis=1)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hlld_adjoint_is_transpose_of_tangent(native, oracle_lib, d):
    _say("case: HLLD dot-product test, direction %d" % d)
    z = golden("rows_hlld.npz")
    vl = np.ascontiguousarray(z["hlld%d_vl" % d])
    vr = np.ascontiguousarray(z["hlld%d_vr" % d])
    rng = np.random.default_rng(100 + d)
    tl = rng.standard_normal(vl.shape)
    tr = rng.standard_normal(vr.shape)
    lam = rng.standard_normal((vl.shape[0], 8))
    out, dout = _tangent(native, d, vl, vr, tl, tr)
    # value part: the reference's HLLD outputs, bit for bit
    assert np.array_equal(out, _oracle_out(oracle_lib, d, vl, vr))
    al, ar = _adjoint(native, d, vl, vr, lam)
    lhs = np.sum(lam * dout, axis=1)
    rhs = np.sum(al * tl, axis=1) + np.sum(ar * tr, axis=1)
    scale = np.sum(np.abs(lam * dout), axis=1) + np.sum(np.abs(al * tl) + np.abs(ar * tr), axis=1)
    err = np.abs(lhs - rhs) / scale
    assert np.all(np.isfinite(err))
    assert err.max() < 1e-12, err.max()


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hlld_tangent_vs_central_differences(native, oracle_lib, d):
    _say("case: HLLD tangent vs oracle central differences, direction %d" % d)
    z = golden("rows_hlld.npz")
    vl = np.ascontiguousarray(z["hlld%d_vl" % d])
    vr = np.ascontiguousarray(z["hlld%d_vr" % d])
    rng = np.random.default_rng(200 + d)
    # perturb the states relative to their size (keeps densities, pressures positive)
    tl = rng.standard_normal(vl.shape) * (np.abs(vl) + 1e-3)
    tr = rng.standard_normal(vr.shape) * (np.abs(vr) + 1e-3)
    h = 1e-7
    _, dout = _tangent(native, d, vl, vr, tl, tr)
    fp = _oracle_out(oracle_lib, d, vl + h * tl, vr + h * tr)
    fm = _oracle_out(oracle_lib, d, vl - h * tl, vr - h * tr)
    fd = (fp - fm) / (2 * h)
    err = np.max(np.abs(fd - dout), axis=1) / (np.max(np.abs(dout), axis=1) + 1e-30)
    # identical left/right states tie in SL = min(slmin, srmin) and
    # SR = max(slmax, srmax): a kink exactly at the point, where a central
    # difference returns the mean of the one-sided slopes and the tangent
    # the branch the forward takes.  Every other row must agree.
    tie = np.all(vl == vr, axis=1)
    _say("  %d tie rows skipped, max err of the rest %.2e" % (tie.sum(), err[~tie].max()))
    assert tie.sum() < 0.1 * len(tie)
    assert err[~tie].max() < 1e-6


# ---------------------------------------------------------------------
# step level
# ---------------------------------------------------------------------

def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _state_tensors(st):
    torch = _torch()
    prim = torch.tensor(st.prim, dtype=torch.float64, device="cuda").contiguous()
    faces = torch.tensor(np.stack([st.bx, st.by, st.bz]), dtype=torch.float64,
                         device="cuda").contiguous()
    return prim, faces


def _rand_dir(rng, shape):
    """Direction on (rho, v, p) and the faces; zero on the cell-centred field."""
    dp = rng.standard_normal((8,) + shape)
    dp[4:7] = 0.0
    df = rng.standard_normal((3,) + shape)
    return dp, df


def _forward(stepper, prim, faces, nsteps):
    stepper.load(prim, faces)
    stepper.run(nsteps)
    return stepper.save()


def _grad_linear(stepper, prim, faces, nsteps, wp, wf):
    """Adjoint of J = <wp, prim_N> + <wf, faces_N> through nsteps steps."""
    torch = _torch()
    states = []
    stepper.load(prim, faces)
    for _ in range(nsteps):
        states.append(stepper.save())
        stepper.run(1)
    lp = torch.tensor(wp, dtype=torch.float64, device="cuda").contiguous()
    lf = torch.tensor(wf, dtype=torch.float64, device="cuda").contiguous()
    for s in reversed(range(nsteps)):
        stepper.load(*states[s])
        stepper.adjoint_step(lp, lf)
    return lp.cpu().numpy(), lf.cpu().numpy()


@pytest.mark.parametrize("nsteps", [1, 3])
def test_step_adjoint_vs_central_differences(native, nsteps):
    from pluto_amd import problems as P
    from pluto_amd.adjoint import Stepper, cfl_dt
    torch = _torch()
    _say("case: step adjoint vs FD, random 12^3, %d steps" % nsteps)
    st = P.random_mhd_3d(12, seed=31, v_rms=0.3, b_rms=0.4)
    shape = st.shape
    dt = cfl_dt(st, 0.3)
    stepper = Stepper(shape, st.spacing(), dt)
    prim, faces = _state_tensors(st)
    rng = np.random.default_rng(7 + nsteps)
    wp = rng.standard_normal((8,) + shape)
    wf = rng.standard_normal((3,) + shape)
    gp, gf = _grad_linear(stepper, prim, faces, nsteps, wp, wf)
    assert np.all(gp[4:7] == 0.0)

    def J(p, f):
        op, of = _forward(stepper, p, f, nsteps)
        return float(np.sum(wp * op.cpu().numpy()) + np.sum(wf * of.cpu().numpy()))

    h = 1e-7
    for trial in range(3):
        dp, df = _rand_dir(rng, shape)
        tdp = torch.tensor(dp, dtype=torch.float64, device="cuda")
        tdf = torch.tensor(df, dtype=torch.float64, device="cuda")
        fd = (J(prim + h * tdp, faces + h * tdf) - J(prim - h * tdp, faces - h * tdf)) / (2 * h)
        ad = float(np.sum(gp * dp) + np.sum(gf * df))
        _say("  dir %d: adjoint %.12e  fd %.12e  rel %.2e" % (trial, ad, fd, abs(ad - fd) / abs(ad)))
        assert abs(ad - fd) <= 1e-6 * abs(ad)
    stepper.close()


def test_taped_adjoint_equals_recomputed(native):
    """pgpu_adjoint_record takes the step pgpu_run takes, bit for bit, and
    the adjoint from its tape equals the adjoint that recomputes the step,
    bit for bit, whatever state the domain holds meanwhile; a taped rollout's
    gradient equals the recomputing rollout's."""
    from pluto_amd import problems as P
    from pluto_amd.adjoint import Stepper, cfl_dt, rollout
    torch = _torch()
    _say("case: taped vs recomputed adjoint, random 16^3")
    st = P.random_mhd_3d(16, seed=51, v_rms=0.3, b_rms=0.4)
    shape = st.shape
    stepper = Stepper(shape, st.spacing(), cfl_dt(st, 0.3))
    prim, faces = _state_tensors(st)
    stepper.load(prim, faces)
    stepper.run(1)
    ref_p, ref_f = stepper.save()
    stepper.load(prim, faces)
    tape = stepper.new_tape()
    stepper.record(tape)
    p1, f1 = stepper.save()
    assert torch.equal(p1, ref_p) and torch.equal(f1, ref_f)
    assert torch.equal(tape[8:11], faces)

    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    lp = torch.randn((8,) + shape, generator=g, dtype=torch.float64, device="cuda")
    lf = torch.randn((3,) + shape, generator=g, dtype=torch.float64, device="cuda")
    lp2, lf2 = lp.clone(), lf.clone()
    stepper.load(prim, faces)
    stepper.adjoint_step(lp, lf)
    stepper.load(ref_p, ref_f)
    stepper.adjoint_step_taped(tape, lp2, lf2)
    assert torch.equal(lp, lp2) and torch.equal(lf, lf2)

    wp = torch.randn((8,) + shape, generator=g, dtype=torch.float64, device="cuda")
    wf = torch.randn((3,) + shape, generator=g, dtype=torch.float64, device="cuda")
    grads = []
    for taped in (True, False):
        p = prim.clone().requires_grad_(True)
        f = faces.clone().requires_grad_(True)
        op, of = rollout(p, f, stepper, 5, chunk=2, taped=taped)
        (torch.sum(op * wp) + torch.sum(of * wf)).backward()
        grads.append((p.grad, f.grad))
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])
    stepper.close()


def test_adjoint_record_refusals(native):
    import ctypes
    from pluto_amd import problems as P
    from pluto_amd.native import lib, check
    torch = _torch()
    st = P.random_mhd_3d(8, seed=3)
    tape = torch.empty((native.PGPU_ADJOINT_TAPE_ARRAYS,) + st.shape, dtype=torch.float64,
                       device="cuda")
    lam = torch.zeros((8,) + st.shape, dtype=torch.float64, device="cuda")
    lamf = torch.zeros((3,) + st.shape, dtype=torch.float64, device="cuda")
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    dom = native.Domain.from_state(st)          # CFL-controlled dt: refused
    with pytest.raises(native.PGPUError):
        check(lib().pgpu_adjoint_record(dom.h, ptr(tape)))
    with pytest.raises(native.PGPUError):
        check(lib().pgpu_adjoint_step_taped(dom.h, ptr(tape), ptr(lam), ptr(lamf)))
    with pytest.raises(native.PGPUError):
        check(lib().pgpu_adjoint_record(dom.h, None))
    dom.close()


def test_adjoint_of_tiled_box_equals_tiled_adjoint(native):
    """A periodic state tiled 3 x 5 x 22 times (z, y, x: 264-cell rows, wider
    than one workgroup, take the x march; y and z march in several segments
    with seams inside the tiles) has as adjoint, for the tiled functional,
    the tiled adjoint of the 12^3 tile (x through LDS rows, one segment per
    line), up to the order of the sums."""
    from pluto_amd import problems as P
    from pluto_amd.adjoint import Stepper, cfl_dt
    torch = _torch()
    _say("case: adjoint of 12^3 tiled to 36 x 60 x 264 vs the tile's, 2 steps")
    st = P.random_mhd_3d(12, seed=41, v_rms=0.3, b_rms=0.4)
    shape = st.shape
    reps = (3, 5, 22)
    dt = cfl_dt(st, 0.3)
    rng = np.random.default_rng(43)
    wp = rng.standard_normal((8,) + shape)
    wf = rng.standard_normal((3,) + shape)
    small = Stepper(shape, st.spacing(), dt)
    gp, gf = _grad_linear(small, *_state_tensors(st), 2, wp, wf)
    small.close()
    big_shape = tuple(s * r for s, r in zip(shape, reps))
    tile = lambda a: np.ascontiguousarray(np.tile(a, (1,) + reps))
    prim = torch.tensor(tile(st.prim), dtype=torch.float64, device="cuda").contiguous()
    faces = torch.tensor(tile(np.stack([st.bx, st.by, st.bz])), dtype=torch.float64,
                         device="cuda").contiguous()
    big = Stepper(big_shape, st.spacing(), dt)
    Gp, Gf = _grad_linear(big, prim, faces, 2, tile(wp), tile(wf))
    big.close()
    for name, a, b in (("prim", Gp, tile(gp)), ("faces", Gf, tile(gf))):
        err = np.abs(a - b).max() / np.abs(b).max()
        _say("  %s: max |difference| / max |adjoint| = %.2e" % (name, err))
        assert err <= 1e-12


def test_one_step_gradient_vs_oracle_differences(native, oracle_lib):
    """The FD reference computed with the oracle's forward step.

    (A state with an exactly uniform field, such as the smooth state's
    pressure, sits on the MC limiter's kink at zero differences: there the
    step has directional derivatives but no gradient, so the state here is
    the non-uniform random one.)"""
    from pluto_amd import problems as P
    from pluto_amd.adjoint import Stepper, cfl_dt
    _say("case: one-step adjoint vs oracle FD, random 10^3")
    st = P.random_mhd_3d(10, seed=3, v_rms=0.3, b_rms=0.4)
    shape = st.shape
    dt = cfl_dt(st, 0.4)
    stepper = Stepper(shape, st.spacing(), dt)
    prim, faces = _state_tensors(st)
    rng = np.random.default_rng(11)
    wp = rng.standard_normal((8,) + shape)
    wf = rng.standard_normal((3,) + shape)
    gp, gf = _grad_linear(stepper, prim, faces, 1, wp, wf)

    def J(dp, df, h):
        s = st.copy()
        s.prim = s.prim + h * dp
        s.bx, s.by, s.bz = (s.bx + h * df[0], s.by + h * df[1], s.bz + h * df[2])
        s.prim[4], s.prim[5], s.prim[6] = P.face_average(s.bx, s.by, s.bz)
        for a in ("prim", "bx", "by", "bz"):
            setattr(s, a, np.ascontiguousarray(getattr(s, a)))
        oracle_lib.rk2_step(s, dt)
        return float(np.sum(wp * s.prim) + np.sum(wf * np.stack([s.bx, s.by, s.bz])))

    h = 1e-7
    for trial in range(2):
        dp, df = _rand_dir(rng, shape)
        fd = (J(dp, df, h) - J(dp, df, -h)) / (2 * h)
        ad = float(np.sum(gp * dp) + np.sum(gf * df))
        _say("  dir %d: adjoint %.12e  oracle fd %.12e" % (trial, ad, fd))
        assert abs(ad - fd) <= 1e-6 * abs(ad)
    stepper.close()


def test_checkpointed_rollout_matches_stored_states(native):
    from pluto_amd import problems as P
    from pluto_amd.adjoint import Stepper, cfl_dt, rollout
    torch = _torch()
    _say("case: rollout gradient, chunk 3 vs chunk 1, 7 steps, 10^3")
    st = P.smooth_mhd_3d(10, seed=4)
    dt = cfl_dt(st, 0.4)
    stepper = Stepper(st.shape, st.spacing(), dt)
    prim0, faces0 = _state_tensors(st)
    target = prim0[0].clone() * 1.01
    grads = []
    for chunk in (1, 3):
        p = prim0.clone().requires_grad_(True)
        f = faces0.clone().requires_grad_(True)
        op, of = rollout(p, f, stepper, 7, chunk=chunk)
        J = torch.sum((op[0] - target) ** 2)
        J.backward()
        grads.append((p.grad.cpu().numpy(), f.grad.cpu().numpy()))
    for a, b in zip(grads[0], grads[1]):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-15 * np.abs(a).max())
    stepper.close()


def test_adjoint_step_is_deterministic(native):
    """Stencil adjoints are gathered per cell, not scattered with atomics:
    two adjoint passes over the same state and seed give the same bits."""
    from pluto_amd import problems as P
    from pluto_amd.adjoint import Stepper, cfl_dt
    _say("case: adjoint determinism, random 12^3, 2 steps")
    st = P.random_mhd_3d(12, seed=21, v_rms=0.3, b_rms=0.4)
    stepper = Stepper(st.shape, st.spacing(), cfl_dt(st, 0.4))
    prim, faces = _state_tensors(st)
    rng = np.random.default_rng(5)
    wp = rng.standard_normal((8,) + st.shape)
    wf = rng.standard_normal((3,) + st.shape)
    a = _grad_linear(stepper, prim, faces, 2, wp, wf)
    b = _grad_linear(stepper, prim, faces, 2, wp, wf)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    stepper.close()


def test_full_gradient_componentwise_vs_oracle(native, oracle_lib):
    """Every component of J^T lam at 6^3 (random state) against central
    differences of the oracle's step, one input component at a time, for
    lam on each output group."""
    from pluto_amd import problems as P
    from pluto_amd.adjoint import Stepper, cfl_dt
    torch = _torch()
    n = 6
    _say("case: full one-step gradient, component by component, random %d^3" % n)
    st = P.random_mhd_3d(n, seed=31, v_rms=0.3, b_rms=0.4)
    shape = st.shape
    dt = cfl_dt(st, 0.4)
    stepper = Stepper(shape, st.spacing(), dt)
    prim, faces = _state_tensors(st)
    x0 = np.concatenate([st.prim, np.stack([st.bx, st.by, st.bz])])

    def out(x):
        s = st.copy()
        s.prim = x[:8].copy()
        s.bx, s.by, s.bz = x[8].copy(), x[9].copy(), x[10].copy()
        s.prim[4], s.prim[5], s.prim[6] = P.face_average(s.bx, s.by, s.bz)
        for a in ("prim", "bx", "by", "bz"):
            setattr(s, a, np.ascontiguousarray(getattr(s, a)))
        oracle_lib.rk2_step(s, dt)
        return np.concatenate([s.prim, np.stack([s.bx, s.by, s.bz])])

    inputs = [0, 1, 2, 3, 7, 8, 9, 10]
    for group in ([0], [1, 2, 3], [7], [8, 9, 10], [4, 5, 6]):
        rng = np.random.default_rng(sum(group))
        lam = np.zeros((11,) + shape)
        for q in group:
            lam[q] = rng.standard_normal(shape)
        lp = torch.tensor(lam[:8], dtype=torch.float64, device="cuda").contiguous()
        lf = torch.tensor(lam[8:], dtype=torch.float64, device="cuda").contiguous()
        stepper.load(prim, faces)
        stepper.adjoint_step(lp, lf)
        g = np.concatenate([lp.cpu().numpy(), lf.cpu().numpy()])
        gfd = np.zeros_like(g)
        for v in inputs:
            for idx in np.ndindex(shape):
                h = 1e-6 * max(1.0, abs(x0[(v,) + idx]))
                xp = x0.copy()
                xp[(v,) + idx] += h
                xm = x0.copy()
                xm[(v,) + idx] -= h
                gfd[(v,) + idx] = np.sum(lam * (out(xp) - out(xm))) / (2 * h)
        err = np.abs(g[inputs] - gfd[inputs]).max() / np.abs(gfd[inputs]).max()
        _say("  lambda on %s: max |adjoint - fd| / max|fd| = %.2e" % (group, err))
        assert err < 1e-7
        assert np.all(g[4:7] == 0.0)
    stepper.close()


def test_control_gradient_componentwise(native):
    """dJ/d(rho0, v0, A0) of J = sum (rho_N - rho_target)^2 through an 8-step
    checkpointed rollout at 16^3 (the 128^3 gradient problem, smaller),
    against central differences of the forward path, one control component
    at a time.  (Directions that perturb every cell at once cross MC-limiter
    branch switches inside +-h at this resolution, which moves a difference
    quotient by O(1e-3); single components rarely do.)"""
    from pluto_amd import problems as P
    from pluto_amd.adjoint import Stepper, cfl_dt, controls_to_state, rollout
    torch = _torch()
    _say("case: control gradient, component by component, smooth 16^3, 8 steps")
    n, nsteps = 16, 8
    rho, v, a = P.smooth_mhd_controls(n, seed=5)
    st = P.smooth_mhd_3d(n, seed=5)
    dt = cfl_dt(st, 0.4)
    d = st.spacing()
    stepper = Stepper(st.shape, d, dt)
    dev = dict(dtype=torch.float64, device="cuda")
    p_t = torch.ones((n, n, n), **dev)
    rho_t = torch.tensor(P.smooth_mhd_3d(n, seed=6).prim[0], **dev)

    def J_of(r, vv, aa):
        prim, faces = controls_to_state(r, vv, aa, p_t, (1.0, 0.0, 0.0), d)
        op, _ = rollout(prim.contiguous(), faces.contiguous(), stepper, nsteps)
        return torch.sum((op[0] - rho_t) ** 2)

    ctl = [torch.tensor(rho, **dev), torch.tensor(v, **dev), torch.tensor(np.stack(a), **dev)]
    leaves = [c.clone().requires_grad_(True) for c in ctl]
    J = J_of(*leaves)
    J.backward()
    g = [x.grad for x in leaves]
    scale = max(float(x.abs().max()) for x in g)
    rng = np.random.default_rng(12)
    h = 1e-7   # at 1e-6 a few components of this 16^3, k <= 4 state straddle a limiter switch
    worst = 0.0
    with torch.no_grad():
        for trial in range(36):
            which = trial % 3
            idx = tuple(int(i) for i in rng.integers(0, n, ctl[which].dim() - 0))
            if ctl[which].dim() == 4:
                idx = (int(rng.integers(0, 3)),) + idx[1:]
            xp = [c.clone() for c in ctl]
            xm = [c.clone() for c in ctl]
            xp[which][idx] += h
            xm[which][idx] -= h
            fd = (J_of(*xp).item() - J_of(*xm).item()) / (2 * h)
            ad = float(g[which][idx])
            worst = max(worst, abs(ad - fd) / scale)
    _say("  36 components: max |adjoint - fd| / max|grad| = %.2e" % worst)
    assert worst < 1e-6
    stepper.close()
