"""Benchmark: RK2 steps of the unsplit MHD stage (PLM + HLLD + CT) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]

Workload: 3D ideal MHD on the periodic unit cube, 256^3, gamma = 5/3, FP64,
seeded random-field state (rho = 1 + 0.1 g, p = 1 + 0.05 g', v and the
edge-centred vector potential with a |k|^-2 amplitude spectrum over
1 <= |k| <= 8, |B|rms = 0.5, div B = 0), PLM (MC) + HLLD + UCT_HLL CT, RK2,
CFL 0.4.  A "step" is one full RK2 step of the whole box; the unit of work is
a zone-update (one zone advanced by one step).

The state is resident in HBM before the timed region; the timed region is
K steps enqueued on the library's stream, bracketed by HIP events on that
stream and by device synchronisation.  Per-kernel durations come from HIP
events recorded around every launch inside the same region.

The same workload is then run with the FP32-flux path (PLM + HLLD in FP32,
conservative accumulation, EMF, CT and ConsToPrim in FP64) and reported
under "fp32_flux" with its speedup and its per-field relative L1/Linf
difference from the FP64 run (computed on the device).  The precision study
at 512^3 over 50 steps is `python bench.py --n 512 --steps 50 --warmup 0`.

Prints one JSON line (rank 0).  cpu_baseline: PLUTO's own serial CPU path
(the reference main loop built from its sources into oracle/_ref) on a
bounded sample of the same workload, on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix), TFLOP/s
ZONE_BYTES = 320.0             # algorithmic bytes per zone-update (5 passes x 64 B)
METRIC = "Mzone-updates/s 256³ HLLD+CT: plain FP64 vs FP32-flux; adjoint/forward cost"

# Algorithmic bytes per cell per launch of each kernel of the current
# pipeline: its unique FP64 inputs and outputs (neighbour re-reads excluded).
# Stage-1 and stage-2 launches alternate, so each entry is their mean.
KERNEL_BYTES = {
    # fused pipeline (default)
    "flux_xy": 0.5 * (208.0 + 240.0),  # V 8, bx by 2, (U 5); U 5, S 4, dv 6, (C_dt 1)
    "flux_z": 0.5 * (200.0 + 192.0),   # V 8, bz 1, U 5, (C_dt 1); U 5, S 2, dv 3
    "emf": 192.0,                      # S 6, dv 9, v 3, faces 3; edges 3
    "ct": 0.5 * (216.0 + 264.0),       # edges 3, faces 3, U 5, (b0 3, V 8); faces 3, V 8, (U 5)
    # modular pipeline
    "sweep_x": 168.0, "sweep_y": 168.0, "sweep_z": 168.0,
    "update": 0.5 * (272.0 + 248.0),
    "c2p": 0.5 * (168.0 + 192.0),
    "step_end": 0.0,
}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_baseline_reference(n=64, steps_a=2, steps_b=50):
    """Time PLUTO's serial CPU path (oracle/_ref/pluto3d) on an n^3 sample.

    Two runs of different length cancel start-up and I/O; rate = zone-updates
    of the extra steps / extra wall time."""
    from oracle import refrun
    from pluto_amd import problems as P
    if not refrun.available(3):
        return None
    st = P.random_mhd_3d(n, seed=2024)
    times = []
    for ns in (steps_a, steps_b):
        t0 = time.perf_counter()
        refrun.run(st, ns, fixed_dt=1e-4)
        times.append(time.perf_counter() - t0)
    # run() executes nsteps + 2 reference steps (it asks for -maxsteps nsteps+1)
    extra = steps_b - steps_a
    dt = times[1] - times[0]
    rate = n ** 3 * extra / dt
    return {"value": rate / 1e6, "unit": "Mzone-updates/s", "cores": 1, "kind": "reference",
            "sample": "%d^3 of the same random-field state, %d RK2 steps timed as the "
                      "difference of a %d- and a %d-step run of PLUTO's serial main loop "
                      "(%.2f s)" % (n, extra, steps_a + 2, steps_b + 2, dt)}


def cpu_port_rate(n=96, steps=6):
    """The oracle (C restatement) with OpenMP over the host threads."""
    from oracle import oracle as O
    from pluto_amd import problems as P
    st = P.random_mhd_3d(n, seed=2024)
    O.rk2_step(st, 1e-4)          # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        O.rk2_step(st, 1e-4)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt / 1e6, "unit": "Mzone-updates/s",
            "cores": O.num_threads(), "kind": "port",
            "sample": "%d^3, %d RK2 steps" % (n, steps)}


def measured_traffic(kernel, n, pipeline):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (profiles/hbm_traffic_fused_256cube.md),
    when this run is the workload they were taken on; otherwise None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_fused_256cube.json")
    if n != 256 or pipeline != "fused" or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["bytes_per_launch"].get(kernel)


# MI355X VALU issue peak: 256 CUs x 4 SIMDs, one wave64 VALU instruction per
# SIMD every 4 cycles (16 lanes per cycle), 2.4 GHz; FP64 vector peak.
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4
FP64_PEAK_FLOPS = 78.6e12


def valu_roofline(kernel, n, pipeline, ms_per_launch):
    """The dominant kernel against the VALU issue rate and FP64 FLOP/s: the
    counts per launch come from the committed counter pass, the launch time
    from this run."""
    m = measured_valu(kernel, n, pipeline)
    if m is None:
        return None
    insts, flop = m
    t = ms_per_launch * 1e-3
    return {"bound": "valu", "kernel": kernel,
            "achieved": insts / t / 1e9, "peak": VALU_PEAK_WAVE_INSTS / 1e9,
            "unit": "G wave-instructions/s", "frac": insts / t / VALU_PEAK_WAVE_INSTS,
            "valu_insts_per_launch": insts,
            "fp64_tflops": flop / t / 1e12, "fp64_peak_tflops": FP64_PEAK_FLOPS / 1e12,
            "fp64_frac": flop / t / FP64_PEAK_FLOPS}


def measured_step_traffic(n, pipeline):
    """HBM bytes of one RK2 step (both stages of every kernel) from the
    committed counter passes, when this run is that workload; else None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_fused_256cube.json")
    if n != 256 or pipeline != "fused" or not os.path.exists(path):
        return None
    with open(path) as f:
        per = json.load(f)["bytes_per_launch"]
    return 2.0 * sum(per.values())


def flop_per_zone_update():
    """The algorithmic FLOP count of one zone-update: the oracle's FP64 step
    compiled with an operation counter (oracle/opcount.py), committed as
    profiles/flop_count_step.json."""
    path = os.path.join(ROOT, "profiles", "flop_count_step.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def host_cpu():
    """CPU model and logical CPU count of this host (the CPU baseline's box)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def measured_valu(kernel, n, pipeline):
    """Wave-level VALU instructions and FP64 FLOP per launch of ``kernel``
    from the committed rocprofv3 SQ counter pass
    (profiles/valu_fused_256cube.md), when this run is that workload."""
    path = os.path.join(ROOT, "profiles", "valu_fused_256cube.json")
    if n != 256 or pipeline != "fused" or not os.path.exists(path):
        return None
    with open(path) as f:
        v = json.load(f)
    if kernel not in v["valu_insts_per_launch"]:
        return None
    return v["valu_insts_per_launch"][kernel], v["f64_flop_per_launch"][kernel]


def checkpoint_study(N, dom, template):
    """PLUTO .dbl output of the domain (device -> host copy, data.NNNN.dbl,
    dbl.out, restart.out) into a temporary directory, then a restart of a
    second domain from it; times both and checks the restarted state."""
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="pgpu_dbl_")
    try:
        t0 = time.perf_counter()
        dom.write_dbl(d, 0)
        t1 = time.perf_counter()
        nbytes = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))
        dom2 = N.Domain(template.shape, dims=template.dims, spacing=template.spacing())
        dom2.set_dt(1.0, 0.4, 1.1)
        t2 = time.perf_counter()
        dom2.read_dbl(d, 0)
        t3 = time.perf_counter()
        a, b = dom.get_state(template), dom2.get_state(template)
        same = all(np.array_equal(getattr(a, f), getattr(b, f)) for f in ("prim", "bx", "by", "bz"))
        same = same and dom.scalars()["time"] == dom2.scalars()["time"]
        dom2.close()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"format": "PLUTO .dbl single_file + dbl.out + restart.out", "bytes": nbytes,
            "write_s": t1 - t0, "restart_s": t3 - t2,
            "write_gbs": nbytes / (t1 - t0) / 1e9, "restart_gbs": nbytes / (t3 - t2) / 1e9,
            "restart_identical": bool(same),
            "note": "host path: device<->host copies over PCIe plus the file system of the run directory"}


def batch_study(npatch=4096, sizes=(16, 24, 32, 48, 64), seeds_per_size=8, steps=5, seed=2024):
    """Batched patches: npatch independent periodic cubes with edges drawn
    uniformly from ``sizes`` (seeded), each holding the random-field state
    of the 256^3 workload generated at its own size (seeds_per_size seeds
    per size, each patch a seeded periodic translate of one of them),
    advanced by the default stacked step (the patches of each edge stacked and
    run through the fused single-domain kernels, the step replayed from a
    hipGraph), and for comparison by the persistent patch kernel (one launch
    per RK stage) at the launch shape the autotuner picks for it.
    Zone-updates count interior zones."""
    import torch
    from pluto_amd import native as N
    from pluto_amd import problems as P
    torch.cuda.empty_cache()
    rng = np.random.default_rng(seed)
    edges = [int(x) for x in rng.choice(sizes, npatch)]
    base = {n: [P.random_mhd_3d(n, seed=seed + 1000 * n + s) for s in range(seeds_per_size)]
            for n in sizes}
    b = N.Batch(edges)
    for p, n in enumerate(edges):
        st = base[n][int(rng.integers(seeds_per_size))]
        sh = tuple(int(x) for x in rng.integers(0, n, 3))
        prim = np.roll(st.prim, sh, axis=(1, 2, 3))
        faces = [np.roll(a, sh, axis=(0, 1, 2)) for a in (st.bx, st.by, st.bz)]
        b.set_patch(p, P.MHDState(prim, *faces, 3, (1.0, 1.0, 1.0)))
    b.set_dt(1e-4, 0.4, 1.1)
    mode, stacked = b.mode()
    b.run(1)
    b.sync()
    ms = b.time_steps(steps) / steps
    # the persistent patch kernel on the same patches, tuned
    b.set_mode(0)
    threads, groups, tune_ms = b.tune(2)
    b.run(1)
    b.sync()
    ms_persistent = b.time_steps(steps) / steps
    b.set_mode(mode)
    zones = sum(n ** 3 for n in edges)
    # stacks: one per edge, split where a field of the stack would reach 4 GiB
    nstacks = sum(-(-edges.count(n) * n ** 3 * 8 // (1 << 32)) if edges.count(n) else 0 for n in sizes)
    out = {
        "workload": "%d periodic patches, edges uniform in %s (seed %d), random-field "
                    "states at each patch's own size, FP64, per-patch CFL 0.4" %
                    (npatch, list(sizes), seed),
        "zones": zones,
        "zones_with_2_ghost_layers": sum((n + 4) ** 3 for n in edges),
        "ms_per_step": ms,
        "value": zones / (ms * 1e-3) / 1e6,
        "unit": "Mzone-updates/s",
        "mode": "stacked" if mode == 1 else "persistent",
        "stacked_patches": stacked,
        "launches_per_step": 4 + 8 * nstacks if stacked == npatch else None,
        "persistent": {"ms_per_step": ms_persistent, "value": zones / (ms_persistent * 1e-3) / 1e6,
                       "launch": {"threads": threads, "groups": groups, "cell_tile": list(b.tile()),
                                  "tuned_ms_per_step": tune_ms},
                       "launches_per_step": 2},
        "device_bytes": b.nbytes(),
    }
    b.close()
    return out


def dropin_study(st, steps=5, dt=1e-4):
    """PLUTO's AdvanceStep through the C-ABI on PLUTO-layout host arrays
    (ARRAY_4D Vc with 2 ghost zones, ArrayBox faces): the rate a PLUTO build
    sees with the library in place of rk_step.o, and where the time goes
    (PGPU_DropinTimes).  The result is checked against the resident domain
    run from the same state at the same dt."""
    import ctypes
    from pluto_amd import native as N
    from pluto_amd import plutoabi as A

    lib = A.load_abi(N.LIB_PATH)

    class Times(ctypes.Structure):
        _fields_ = [(k, ctypes.c_double) for k in
                    ("pack", "h2d", "step", "d2h", "unpack", "total")] + [("calls", ctypes.c_longlong)]

    lib.PGPU_DropinTimes.argtypes = [ctypes.POINTER(Times), ctypes.c_int]
    box = A.PlutoBox(st)
    hlld = ctypes.cast(lib.HLLD_Solver, ctypes.c_void_p).value
    lib.SetSolver(b"hlld")
    ctypes.c_double.in_dll(lib, "g_dt").value = dt

    def step():
        ctypes.c_double.in_dll(lib, "g_maxMach").value = 0.0
        box.dts.invDt_hyp = 0.0
        if lib.AdvanceStep(ctypes.byref(box.data), hlld, ctypes.byref(box.dts),
                           ctypes.byref(box.grid)) != 0:
            raise RuntimeError("AdvanceStep failed")

    step()                                  # allocations: device box, pinned staging
    t = Times()
    lib.PGPU_DropinTimes(ctypes.byref(t), 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    wall = time.perf_counter() - t0
    lib.PGPU_DropinTimes(ctypes.byref(t), 1)
    out = box.get_state()
    lib.pgpu_finalize()
    dom = N.Domain.from_state(st)
    dom.set_dt(dt, fixed=True)
    dom.run(steps + 1)
    ref = dom.get_state(st)
    dom.close()
    same = all(np.array_equal(a, b) for a, b in
               ((out.prim, ref.prim), (out.bx, ref.bx), (out.by, ref.by), (out.bz, ref.bz)))
    zones = st.prim[0].size
    per = {k: getattr(t, k) / t.calls * 1e3 for k in ("pack", "h2d", "step", "d2h", "unpack",
                                                      "total")}
    return {
        "entry_point": "AdvanceStep (include/pluto_abi.h), periodic box, HLLD, UCT_HLL, "
                       "fixed dt %g" % dt,
        "steps": steps,
        "value": zones * steps / wall / 1e6,
        "unit": "Mzone-updates/s",
        "ms_per_step": wall / steps * 1e3,
        "ms_per_call": per,
        "host_bytes_per_call": 2 * 11 * 8 * zones,
        "pcie_GBps": {"h2d": 11 * 8 * zones / (per["h2d"] * 1e-3) / 1e9,
                      "d2h": 11 * 8 * zones / (per["d2h"] * 1e-3) / 1e9},
        "matches_resident_domain": bool(same),
        "note": "pack/unpack are host copies between PLUTO's arrays and pinned buffers "
                "(pack overlaps h2d, unpack overlaps d2h); h2d/step/d2h are device-event "
                "intervals",
    }


def adjoint_study(n=128, nsteps=64, ndirs=32, nplain=8, ncomp=12, h=1e-6, h1=1e-5, seed=5):
    """Gradient problem on the smooth 128^3 state: J = sum (rho_N - rho_t)^2
    after nsteps fixed-dt RK2 steps, dJ/d(rho0, v0, A0) by the HIP discrete
    adjoint with sqrt(N) checkpointing.  Reports the cost of an adjoint step
    and of the whole gradient against forward steps, and the central
    difference checks: ndirs random directions with the branch outcomes of
    the run at x imposed (the piece the adjoint differentiates), nplain of
    them as plain differences screened by the branch signature, and ncomp
    single control components."""
    import torch
    from pluto_amd import native as N
    from pluto_amd import problems as P
    from pluto_amd.adjoint import Stepper, cfl_dt, controls_to_state, rollout

    rho, v, a = P.smooth_mhd_controls(n, seed=seed)
    st = P.smooth_mhd_3d(n, seed=seed)
    dt = cfl_dt(st, 0.4)
    d = st.spacing()
    stepper = Stepper(st.shape, d, dt)
    dev = dict(dtype=torch.float64, device="cuda")
    p_t = torch.ones((n, n, n), **dev)
    rho_t = torch.tensor(P.smooth_mhd_3d(n, seed=seed + 1).prim[0], **dev)
    r = torch.tensor(rho, **dev)
    vv = torch.tensor(v, **dev)
    aa = torch.tensor(np.stack(a), **dev)

    def J_of(r_, v_, a_):
        prim, faces = controls_to_state(r_, v_, a_, p_t, (1.0, 0.0, 0.0), d)
        op, _ = rollout(prim.contiguous(), faces.contiguous(), stepper, nsteps)
        return torch.sum((op[0] - rho_t) ** 2)

    # forward step and adjoint step, each timed on the library stream
    prim0, faces0 = controls_to_state(r, vv, aa, p_t, (1.0, 0.0, 0.0), d)
    prim0, faces0 = prim0.contiguous(), faces0.contiguous()
    stepper.load(prim0, faces0)
    stepper.run(2)
    stepper.dom.sync()
    fwd_ms = stepper.dom.time_steps(20) / 20
    stepper.load(prim0, faces0)
    lp = torch.randn((8, n, n, n), **dev)
    lf = torch.randn((3, n, n, n), **dev)
    stepper.adjoint_step(lp, lf)          # allocates the adjoint arrays
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 10
    for _ in range(reps):
        stepper.adjoint_step(lp, lf)
    torch.cuda.synchronize()
    adj_ms = (time.perf_counter() - t0) * 1e3 / reps
    # the same step stored (pgpu_adjoint_record) and reversed from its tape
    tape = stepper.new_tape()
    stepper.load(prim0, faces0)
    stepper.record(tape)
    stepper.adjoint_step_taped(tape, lp, lf)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        stepper.load(prim0, faces0)
        stepper.record(tape)
    stepper.dom.sync()
    rec_ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        stepper.adjoint_step_taped(tape, lp, lf)
    torch.cuda.synchronize()
    adj_taped_ms = (time.perf_counter() - t0) * 1e3 / reps
    del tape

    # whole gradient (forward with checkpoints + backward) vs plain forward
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stepper.load(prim0, faces0)
    stepper.run(nsteps)
    stepper.dom.sync()
    fwd_total = time.perf_counter() - t0
    rg = r.clone().requires_grad_(True)
    vg = vv.clone().requires_grad_(True)
    ag = aa.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = J_of(rg, vg, ag)
    J.backward()
    torch.cuda.synchronize()
    grad_total = time.perf_counter() - t0
    g = [rg.grad, vg.grad, ag.grad]

    ctl = [r, vv, aa]
    dom = stepper.dom

    def final_rho(x, mode):
        """rho after nsteps from controls x: a plain run with its branch
        signature (mode "sig") or a run replaying the branch tape of the run
        at the unperturbed controls ("replay")."""
        pr, fa = controls_to_state(*x, p_t, (1.0, 0.0, 0.0), d)
        stepper.load(pr.contiguous(), fa.contiguous())
        if mode == "sig":
            dom.set_branch_signature(True)
        else:
            dom.set_branch_tape(N.TAPE_REPLAY)
        stepper.run(nsteps)
        out, _ = stepper.save()
        sig = dom.branch_signature() if mode == "sig" else None
        dom.set_branch_signature(False)
        return out[0].clone(), sig

    def diff(x, dirs, hh, mode):
        """J(x + hh d) - J(x - hh d) summed cell by cell, and whether both
        runs kept the signature of the run at x."""
        rp, sp = final_rho([c + hh * e for c, e in zip(x, dirs)], mode)
        rm, sm = final_rho([c - hh * e for c, e in zip(x, dirs)], mode)
        return float(torch.sum((rp - rm) * (rp + rm - 2.0 * rho_t))), (sp, sm)

    def direction(t):
        rng_t = np.random.default_rng([1234, t])
        return [torch.tensor(rng_t.standard_normal(x.shape), **dev) for x in ctl]

    plain, kept, taped, two_point, comp = [], 0, [], [], []
    with torch.no_grad():
        # plain differences, screened by the branch signature
        _, sig0 = final_rho(ctl, "sig")
        for t in range(nplain):
            dirs = direction(t)
            ad = sum(float(torch.sum(gi * di)) for gi, di in zip(g, dirs))
            dj, (sp, sm) = diff(ctl, dirs, h, "sig")
            kept += (sp == sig0 and sm == sig0)
            plain.append(abs(dj / (2 * h) - ad) / abs(ad))
        # the branch tape of the run at x, then taped differences
        pr, fa = controls_to_state(*ctl, p_t, (1.0, 0.0, 0.0), d)
        stepper.load(pr.contiguous(), fa.contiguous())
        dom.set_branch_tape(N.TAPE_RECORD, nsteps)
        stepper.run(nsteps)
        stepper.dom.sync()
        for t in range(ndirs):
            dirs = direction(t)
            ad = sum(float(torch.sum(gi * di)) for gi, di in zip(g, dirs))
            q1 = diff(ctl, dirs, h, "replay")[0] / (2 * h)
            q2 = diff(ctl, dirs, 2 * h, "replay")[0] / (4 * h)
            taped.append(abs((4.0 * q1 - q2) / 3.0 - ad) / abs(ad))
            two_point.append(abs(q1 - ad) / abs(ad))
        # single control components (one cell of rho, v or A), taped, step h1
        rng = np.random.default_rng(1234)
        for t in range(ncomp):
            which = t % 3
            idx = tuple(int(i) for i in rng.integers(0, n, 3))
            if ctl[which].dim() == 4:
                idx = (int(rng.integers(0, 3)),) + idx
            dirs = [torch.zeros_like(c) for c in ctl]
            dirs[which][idx] = 1.0
            fd = diff(ctl, dirs, h1, "replay")[0] / (2 * h1)
            comp.append(abs(float(g[which][idx]) - fd) / abs(float(g[which][idx])))
        dom.set_branch_tape(N.TAPE_OFF)
    stepper.close()
    plain, taped, two_point, comp = (np.array(x) for x in (plain, taped, two_point, comp))
    return {
        "workload": "smooth 3D MHD %d^3 (rho=1, p=1, B=(1,0,0) + |k|<=4 perturbations of "
                    "rms 0.05 on rho, v, A; seed %d), J = sum (rho_N - rho_target)^2 after "
                    "%d RK2 steps at fixed dt %.6g, rho_target from seed %d" %
                    (n, seed, nsteps, dt, seed + 1),
        "forward_step_ms": fwd_ms,
        "adjoint_step_ms": adj_ms,
        "adjoint_over_forward": adj_ms / fwd_ms,
        "taped": {
            "note": "the forward step stored by pgpu_adjoint_record (state load + step + "
                    "tape writes) and its reverse from the tape (pgpu_adjoint_step_taped); "
                    "the gradient below runs this way",
            "record_ms": rec_ms,
            "adjoint_step_ms": adj_taped_ms,
            "adjoint_over_forward": adj_taped_ms / fwd_ms,
            "record_plus_adjoint_over_forward": (rec_ms + adj_taped_ms) / fwd_ms,
        },
        "gradient_s": grad_total,
        "forward_s": fwd_total,
        "gradient_over_forward": grad_total / fwd_total,
        "checkpoint_chunk": int(np.ceil(np.sqrt(nsteps))),
        "fd_check": {
            "h": h,
            "random_directions_taped": {
                "count": ndirs, "max_rel_err": float(taped.max()),
                "median_rel_err": float(np.median(taped)),
                "within_1e-6": int(np.sum(taped <= 1e-6)),
                "two_point_max_rel_err": float(two_point.max()),
                "note": "N(0,1) per cell of rho, v and A; differences of the objective "
                        "with the branch outcomes of the run at x imposed (branch tape), "
                        "fourth-order central stencil at h and 2h; the two-point "
                        "quotient's error is its h^2 truncation term"},
            "random_directions_plain": {
                "count": nplain, "signature_kept": kept,
                "max_rel_err": float(plain.max()), "median_rel_err": float(np.median(plain)),
                "note": "plain two-point differences; the branch signature shows every "
                        "such direction crossing limiter/HLLD switches inside +-h"},
            "single_components_taped": {
                "count": ncomp, "h": h1, "max_rel_err": float(comp.max()),
                "median_rel_err": float(np.median(comp)),
                "within_1e-6": int(np.sum(comp <= 1e-6))},
        },
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--pipeline", choices=["fused", "modular"], default="fused")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-flux run")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint study")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-patch study")
    ap.add_argument("--no-solvers", action="store_true",
                    help="skip the HLL, HLLC, TVDLF and Roe solver and the other EMF-average timings")
    ap.add_argument("--no-checkpoint", action="store_true",
                    help="skip the .dbl checkpoint write/read timing")
    ap.add_argument("--no-fast", action="store_true",
                    help="skip the tolerance-mode FP64 run")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the AdvanceStep (C-ABI, host arrays) timing")
    ap.add_argument("--batch-patches", type=int, default=4096)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one plain 512 MiB device-to-device copy (a known "
                         "byte count for scaling hardware counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the workload is one box on one "
              "MI355X); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch  # device plumbing only
    from pluto_amd import native as N
    from pluto_amd import problems as P

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(3)
    torch.cuda.set_device(0)
    if args.copy_calibration:
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log("calibration copy: 512 MiB read, 512 MiB written")
    n = args.n
    log("generating %d^3 random-field state" % n)
    st = P.random_mhd_3d(n, seed=args.seed)
    pipeline = N.PIPELINE_FUSED if args.pipeline == "fused" else N.PIPELINE_MODULAR

    def timed(precision, solver=N.SOLVER_HLLD, emf_average=N.EMF_UCT_HLL, bounds=None):
        """warmup + K timed steps of a fresh domain; returns (domain, ms/step, kernels)."""
        dom = N.Domain.from_state(st, pipeline=pipeline, precision=precision, solver=solver,
                                  emf_average=emf_average, bounds=bounds)
        dom.set_dt(1e-4, 0.4, 1.1)
        log("warmup %d steps" % args.warmup)
        dom.run(args.warmup)
        dom.sync()
        # timed region: K steps, kernel events on, bracketed by stream events
        dom.timing(True)
        torch.cuda.synchronize()
        log("timing %d steps" % args.steps)
        ms = dom.time_steps(args.steps)
        torch.cuda.synchronize()
        return dom, ms / args.steps, dom.timing_read()

    log("FP64")
    dom, ms_step, kt = timed(N.PGPU_PREC_FP64)
    sc = dom.scalars()
    zups = n ** 3 / (ms_step * 1e-3)

    # dominant kernel by total time
    dom_name = max(kt, key=lambda k: kt[k][0])
    k_ms, k_cnt = kt[dom_name]
    k_avg_ms = k_ms / max(k_cnt, 1)
    k_bytes = KERNEL_BYTES[dom_name] * n ** 3
    k_gbs = k_bytes / (k_avg_ms * 1e-3) / 1e9
    kernels = {k: {"ms_per_launch": v[0] / max(v[1], 1), "launches": v[1]}
               for k, v in kt.items() if v[1]}

    ckpt = None
    if not args.no_checkpoint:
        log("checkpoint (.dbl write and restart)")
        ckpt = checkpoint_study(N, dom, st)

    solvers = None
    if not args.no_solvers:
        solvers = {}
        for name, code in (("hll", N.SOLVER_HLL), ("hllc", N.SOLVER_HLLC),
                           ("tvdlf", N.SOLVER_TVDLF), ("roe", N.SOLVER_ROE)):
            log("%s solver (FP64)" % name)
            dh, msh, kth = timed(N.PGPU_PREC_FP64, code)
            solvers[name] = {"value": n ** 3 / (msh * 1e-3) / 1e6, "unit": "Mzone-updates/s",
                             "ms_per_step": msh, "pluto_solver": name,
                             "kernels": {k: {"ms_per_launch": v[0] / max(v[1], 1), "launches": v[1]}
                                         for k, v in kth.items() if v[1]}}
            dh.close()
        if args.pipeline == "fused":
            solvers["emf_averages"] = {}
            for name in ("arithmetic", "uct0", "uct_contact"):
                log("%s EMF average (HLLD, FP64)" % name)
                dh, msh, kth = timed(N.PGPU_PREC_FP64, N.SOLVER_HLLD, N.EMF_AVERAGES[name])
                solvers["emf_averages"][name] = {
                    "value": n ** 3 / (msh * 1e-3) / 1e6, "unit": "Mzone-updates/s",
                    "ms_per_step": msh, "ct_emf_average": name.upper(),
                    "kernels": {k: {"ms_per_launch": v[0] / max(v[1], 1), "launches": v[1]}
                                for k, v in kth.items() if v[1]}}
                dh.close()
            # non-periodic sides: the bounded box (two ghost zones per side)
            solvers["boundaries"] = {}
            for name, bounds in (("outflow", ("outflow",) * 6),
                                 ("reflective", ("reflective",) * 6),
                                 ("mixed", ("reflective", "outflow", "periodic", "periodic",
                                            "eqtsymmetric", "axisymmetric"))):
                log("%s boundaries (HLLD, FP64)" % name)
                dh, msh, kth = timed(N.PGPU_PREC_FP64, bounds=bounds)
                solvers["boundaries"][name] = {
                    "value": n ** 3 / (msh * 1e-3) / 1e6, "unit": "Mzone-updates/s",
                    "ms_per_step": msh, "pluto_boundary": list(bounds),
                    "kernels": {k: {"ms_per_launch": v[0] / max(v[1], 1), "launches": v[1]}
                                for k, v in kth.items() if v[1]}}
                dh.close()

    fp32 = None
    if args.pipeline == "fused" and not args.no_fp32:
        log("FP32 flux")
        d32, ms32, kt32 = timed(N.PGPU_PREC_FP32_FLUX)
        err = d32.error_vs(dom)
        fp32 = {
            "value": n ** 3 / (ms32 * 1e-3) / 1e6,
            "unit": "Mzone-updates/s",
            "ms_per_step": ms32,
            "speedup_vs_fp64": ms_step / ms32,
            "kernels": {k: {"ms_per_launch": v[0] / max(v[1], 1), "launches": v[1]}
                        for k, v in kt32.items() if v[1]},
            "error_vs_fp64": {
                "after_steps": args.warmup + args.steps,
                "note": "both runs from the same state with their own CFL dt control",
                "l1_rel": {k: v[0] for k, v in err.items()},
                "linf_rel": {k: v[1] for k, v in err.items()},
            },
            "dt_ratio": d32.scalars()["dt"] / sc["dt"],
        }
        d32.close()

    fast = None
    if args.pipeline == "fused" and not args.no_fast:
        log("FP64 tolerance mode (contracted, Newton division/sqrt)")
        df, msf, ktf = timed(N.PGPU_PREC_FP64_FAST)
        err = df.error_vs(dom)
        fast = {
            "value": n ** 3 / (msf * 1e-3) / 1e6,
            "unit": "Mzone-updates/s",
            "ms_per_step": msf,
            "speedup_vs_fp64": ms_step / msf,
            "kernels": {k: {"ms_per_launch": v[0] / max(v[1], 1), "launches": v[1]}
                        for k, v in ktf.items() if v[1]},
            "error_vs_fp64": {
                "after_steps": args.warmup + args.steps,
                "note": "both runs from the same state with their own CFL dt control; "
                        "bit-exact FP64 is the reference",
                "l1_rel": {k: v[0] for k, v in err.items()},
                "linf_rel": {k: v[1] for k, v in err.items()},
            },
        }
        df.close()

    dropin = None
    if not args.no_dropin:
        log("drop-in AdvanceStep through the C-ABI (%d^3 host arrays)" % n)
        dropin = dropin_study(st)

    adj = None
    if not args.no_adjoint:
        log("adjoint study (128^3, 64 steps)")
        dom.close()
        adj = adjoint_study()

    batch = None
    if not args.no_batch:
        log("batched patches (%d)" % args.batch_patches)
        dom.close()
        batch = batch_study(args.batch_patches)

    cpu = None
    cpu_port = None
    if not args.no_cpu:
        log("cpu baseline (reference serial main loop)")
        try:
            cpu = cpu_baseline_reference()
        except Exception as e:  # the reference build may be absent
            log("reference baseline unavailable: %s" % e)
        try:
            cpu_port = cpu_port_rate()
        except Exception as e:
            log("port baseline unavailable: %s" % e)
        if cpu is None:
            cpu = cpu_port

    step_traffic = measured_step_traffic(n, args.pipeline)
    fc = flop_per_zone_update()
    compute_roofline = None
    if fc is not None:
        fpz = fc["flop_per_zone_update"]
        compute_roofline = {
            "bound": "fp64",
            "flop_per_zone_update": fpz,
            "source": "profiles/flop_count_step.json (" + fc["command"] + ")",
            "achieved": zups * fpz / 1e12,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": zups * fpz / 1e12 / FP64_PEAK_TFLOPS,
        }
    if cpu is not None:
        cpu = dict(cpu, host=host_cpu())
    if cpu_port is not None:
        cpu_port = dict(cpu_port, host=host_cpu())

    out = {
        "metric": METRIC,
        "value": zups / 1e6,
        "unit": "Mzone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "3D ideal MHD %d^3 periodic unit cube, seeded random-field state "
                        "(seed %d, 1<=|k|<=8, |k|^-2 amplitude, |B|rms=0.5), "
                        "PLM(MC)+HLLD+CT(UCT_HLL), RK2, CFL 0.4, gamma 5/3" % (n, args.seed),
            "n": n, "zones": n ** 3, "precision": "fp64", "parallelism": "single GPU",
            "pipeline": args.pipeline,
        },
        "roofline": {
            "bound": "hbm",
            "basis": "one RK2 step of the whole box: 320 B per zone-update (the 8-double "
                     "state read at stage 1, written, read and blended with the step-start "
                     "state at stage 2, written: 5 x 64 B), over the step time",
            "achieved": zups * ZONE_BYTES / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": zups * ZONE_BYTES / 1e9 / HBM_PEAK_GBS,
            "bytes_per_zone_update": ZONE_BYTES,
            "traffic": step_traffic,
            "traffic_per_zone_update": (step_traffic / n ** 3) if step_traffic else None,
        },
        "compute_roofline": compute_roofline,
        "kernel_roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": k_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": k_gbs / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom_name, n, args.pipeline),
            "bytes_per_cell": KERNEL_BYTES[dom_name],
            "ms_per_launch": k_avg_ms,
        },
        "valu_roofline": valu_roofline(dom_name, n, args.pipeline, k_avg_ms),
        "kernels": kernels,
        "fp32_flux": fp32,
        "fp64_fast": fast,
        "adjoint": adj,
        "batch": batch,
        "checkpoint": ckpt,
        "dropin": dropin,
        "solvers": solvers,
        "dt_last": sc["dt"],
        "cpu_baseline": cpu,
        "cpu_port": cpu_port,
    }
    dom.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
