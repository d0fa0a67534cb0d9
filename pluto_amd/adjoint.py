"""Gradients through RK2 steps of the MHD stage.

The HIP discrete adjoint of one step (``pgpu_adjoint_step``) is wrapped in a
``torch.autograd.Function`` that checkpoints every ``chunk`` steps on the
forward pass (default chunk = ceil(sqrt(nsteps))) and, on the backward pass,
retakes each segment's steps from its checkpoint, storing what each step's
adjoint reads (``pgpu_adjoint_record``), before running the adjoint steps
backwards through it from those tapes (``pgpu_adjoint_step_taped``).
Memory: about nsteps/chunk states and chunk tapes of 60 zone arrays.  PyTorch holds the device buffers and composes objectives and the
control-to-state map; every forward and adjoint step runs in
libpluto_amd.so (there is no other implementation to fall back to).

The time step is fixed (set from the initial state's CFL condition), so dt
is not a function of the controls.
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import native as N

F64 = torch.float64


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    assert t.is_cuda and t.dtype == F64 and t.is_contiguous()
    return ctypes.c_void_p(t.data_ptr())


class Stepper:
    """A resident 3D periodic domain advanced with a fixed dt."""

    def __init__(self, shape, spacing, dt, gamma=5.0 / 3.0, pipeline=N.PIPELINE_FUSED):
        self.shape = tuple(shape)
        self.dom = N.Domain(self.shape, 3, spacing, gamma, pipeline=pipeline)
        self.dt = float(dt)
        self.dom.set_dt(self.dt, fixed=True)

    def load(self, prim: torch.Tensor, faces: torch.Tensor):
        torch.cuda.current_stream().synchronize()
        N.check(N.lib().pgpu_set_state_device(self.dom.h, _ptr(prim), _ptr(faces)))

    def save(self):
        prim = torch.empty((8,) + self.shape, dtype=F64, device="cuda")
        faces = torch.empty((3,) + self.shape, dtype=F64, device="cuda")
        torch.cuda.current_stream().synchronize()
        N.check(N.lib().pgpu_get_state_device(self.dom.h, _ptr(prim), _ptr(faces)))
        self.dom.sync()
        return prim, faces

    def run(self, nsteps: int):
        self.dom.run(nsteps)

    def adjoint_step(self, lam_prim: torch.Tensor, lam_faces: torch.Tensor):
        torch.cuda.current_stream().synchronize()
        N.check(N.lib().pgpu_adjoint_step(self.dom.h, _ptr(lam_prim), _ptr(lam_faces)))
        self.dom.sync()

    def new_tape(self) -> torch.Tensor:
        """Device buffer for one step's tape (PGPU_ADJOINT_TAPE_ARRAYS zone arrays)."""
        return torch.empty((N.PGPU_ADJOINT_TAPE_ARRAYS,) + self.shape, dtype=F64, device="cuda")

    def record(self, tape: torch.Tensor):
        """One step, storing what its adjoint reads into tape."""
        torch.cuda.current_stream().synchronize()
        N.check(N.lib().pgpu_adjoint_record(self.dom.h, _ptr(tape)))

    def adjoint_step_taped(self, tape: torch.Tensor, lam_prim: torch.Tensor,
                           lam_faces: torch.Tensor):
        """adjoint_step for the step tape holds (no recomputation)."""
        torch.cuda.current_stream().synchronize()
        N.check(N.lib().pgpu_adjoint_step_taped(self.dom.h, _ptr(tape), _ptr(lam_prim),
                                                _ptr(lam_faces)))
        self.dom.sync()

    def close(self):
        self.dom.close()


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, prim, faces, stepper: Stepper, nsteps: int, chunk: int, taped: bool):
        prim = prim.detach().contiguous()
        faces = faces.detach().contiguous()
        stepper.load(prim, faces)
        ckpts = []
        for s in range(nsteps):
            if s % chunk == 0:
                ckpts.append(stepper.save())
            stepper.run(1)
        out = stepper.save()
        ctx.stepper, ctx.nsteps, ctx.chunk, ctx.ckpts = stepper, nsteps, chunk, ckpts
        ctx.taped = taped
        return out

    @staticmethod
    def backward(ctx, g_prim, g_faces):
        st, nsteps, chunk = ctx.stepper, ctx.nsteps, ctx.chunk
        lam_p = (g_prim if g_prim is not None else
                 torch.zeros((8,) + st.shape, dtype=F64, device="cuda")).detach().clone()
        lam_f = (g_faces if g_faces is not None else
                 torch.zeros((3,) + st.shape, dtype=F64, device="cuda")).detach().clone()
        lam_p = lam_p.contiguous()
        lam_f = lam_f.contiguous()
        nseg = (nsteps + chunk - 1) // chunk
        tapes = None
        if ctx.taped:
            try:
                tapes = [st.new_tape() for _ in range(min(chunk, nsteps))]
            except torch.cuda.OutOfMemoryError:
                tapes = None
        for seg in reversed(range(nseg)):
            s0, s1 = seg * chunk, min((seg + 1) * chunk, nsteps)
            st.load(*ctx.ckpts[seg])
            if tapes is not None:
                # each step stored as it is taken, then reversed from the tapes
                for s in range(s0, s1):
                    st.record(tapes[s - s0])
                for s in reversed(range(s0, s1)):
                    st.adjoint_step_taped(tapes[s - s0], lam_p, lam_f)
                continue
            states = [ctx.ckpts[seg]]
            for _ in range(s0, s1 - 1):
                st.run(1)
                states.append(st.save())
            for s in reversed(range(s0, s1)):
                st.load(*states[s - s0])
                st.adjoint_step(lam_p, lam_f)
            del states
        ctx.ckpts = None
        del tapes
        return lam_p, lam_f, None, None, None, None


def rollout(prim: torch.Tensor, faces: torch.Tensor, stepper: Stepper, nsteps: int,
            chunk: int | None = None, taped: bool = True):
    """nsteps RK2 steps of (prim [8,nz,ny,nx], faces [3,nz,ny,nx]), differentiable.

    Returns the final (prim, faces).  The cell-centred field of the input
    prim is ignored (recomputed from the faces) and gets a zero gradient.
    taped: the backward pass stores each segment's steps as it retakes them
    (chunk tapes of PGPU_ADJOINT_TAPE_ARRAYS zone arrays) and runs the
    adjoint steps from the tapes; without it (or when the tapes do not fit)
    every adjoint step recomputes its forward step."""
    if chunk is None:
        chunk = max(1, int(math.ceil(math.sqrt(nsteps))))
    return _Rollout.apply(prim, faces, stepper, int(nsteps), int(chunk), bool(taped))


# ---------------------------------------------------------------------
# control-to-state map of the smooth-state gradient problem
# ---------------------------------------------------------------------

def curl_edges(ax, ay, az, d):
    """Faces from the edge-centred vector potential (vec_pot_diff.c:139-169),
    differentiable torch version of problems.curl_edges."""
    dx, dy, dz = d
    bx = (az - torch.roll(az, 1, 1)) / dy - (ay - torch.roll(ay, 1, 0)) / dz
    by = -(az - torch.roll(az, 1, 2)) / dx + (ax - torch.roll(ax, 1, 0)) / dz
    bz = (ay - torch.roll(ay, 1, 2)) / dx - (ax - torch.roll(ax, 1, 1)) / dy
    return torch.stack([bx, by, bz])


def controls_to_state(rho, v, a, p, b0, d):
    """prim [8,...] and faces [3,...] from rho, v [3,...], A [3,...] (edges),
    pressure p and the uniform field b0 = (bx0, by0, bz0)."""
    zero = torch.zeros_like(rho)
    prim = torch.stack([rho, v[0], v[1], v[2], zero, zero, zero, p])
    faces = curl_edges(a[0], a[1], a[2], d)
    faces = faces + torch.tensor(b0, dtype=F64, device=faces.device).view(3, 1, 1, 1)
    return prim, faces


def cfl_dt(st, cfl=0.4, gamma=5.0 / 3.0):
    """CFL dt of a state (max over cells of sum_d (|v_d| + c_f,d)/dx_d, /3),
    the fixed step of the gradient problem."""
    import numpy as np
    v = st.prim
    gp = gamma * v[7]
    b2 = v[4] ** 2 + v[5] ** 2 + v[6] ** 2
    inv = np.zeros_like(v[0])
    for dd, h in enumerate(st.spacing()):
        bn2 = v[4 + dd] ** 2
        cf2 = 0.5 * (gp + b2 + np.sqrt((gp + b2) ** 2 - 4.0 * gp * bn2)) / v[0]
        inv = inv + (np.abs(v[1 + dd]) + np.sqrt(cf2)) / h
    return cfl / (inv.max() / 3.0)
