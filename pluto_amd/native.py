"""ctypes binding of libpluto_amd.so and a small host-side Domain class.

The product path is the HIP library only: if ``libpluto_amd.so`` is missing
or cannot be loaded, :func:`lib` raises -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from .problems import MHDState

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpluto_amd.so")

PGPU_PREC_FP64 = 0
PGPU_PREC_FP32_FLUX = 1
PGPU_PREC_FP64_FAST = 2   # FP64 flux with FMA contraction and Newton div/sqrt (tolerance mode)
PGPU_ADJOINT_TAPE_ARRAYS = 60   # arrays of n doubles per step tape (pluto_amd.h)
KERNEL_IDS = ("sweep_x", "sweep_y", "sweep_z", "update", "emf", "ct", "c2p", "step_end",
              "flux_xy", "flux_z", "boundary")
FIELDS = ("rho", "mx", "my", "mz", "E", "bx", "by", "bz")
TAPE_OFF, TAPE_RECORD, TAPE_REPLAY = 0, 1, 2   # PGPU_TAPE_*
PIPELINE_FUSED = 0
PIPELINE_MODULAR = 1
SOLVER_HLLD = 0          # PGPU_SOLVER_HLLD: HLLD_Solver (default)
SOLVER_HLL = 1           # PGPU_SOLVER_HLL: HLL_Solver
SOLVER_HLLC = 2          # PGPU_SOLVER_HLLC: HLLC_Solver
SOLVER_TVDLF = 3         # PGPU_SOLVER_TVDLF: LF_Solver ("tvdlf")
SOLVER_ROE = 4           # PGPU_SOLVER_ROE: Roe_Solver ("roe")
# PGPU_EMF_*: PLUTO's CT_EMF_AVERAGE labels (ct.h:18-21)
EMF_ARITHMETIC = 1
EMF_UCT0 = 2
EMF_UCT_CONTACT = 3
EMF_UCT_HLL = 4
EMF_AVERAGES = {"arithmetic": EMF_ARITHMETIC, "uct0": EMF_UCT0, "uct_contact": EMF_UCT_CONTACT,
                "uct_hll": EMF_UCT_HLL}
# PGPU_BOUND_*: PLUTO's boundary types (pluto.h:146-150) by their pluto.ini words
BOUNDS = {"outflow": 1, "reflective": 2, "axisymmetric": 3, "eqtsymmetric": 4, "periodic": 5,
          "userdef": 7}


class Slab(ctypes.Structure):
    """pgpu_slab: the slab of a USERDEF side handed to its hook."""
    _fields_ = [("lo", ctypes.c_int * 3), ("hi", ctypes.c_int * 3), ("ld", ctypes.c_int * 3),
                ("beg", ctypes.c_int * 3), ("end", ctypes.c_int * 3), ("tot", ctypes.c_int * 3),
                ("vc", ctypes.POINTER(ctypes.c_double) * 8),
                ("vs", ctypes.POINTER(ctypes.c_double) * 3)]

    def arrays(self):
        """numpy views ([k][j][i] over lo..hi) of the 8 cell arrays and the
        3 face arrays."""
        shape = (self.ld[2], self.ld[1], self.ld[0])
        vc = [np.ctypeslib.as_array(self.vc[q], shape) for q in range(8)]
        vs = [np.ctypeslib.as_array(self.vs[a], shape) for a in range(3)]
        return vc, vs


USERDEF_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                              ctypes.POINTER(Slab))

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_lib = None


class PGPUError(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PGPUError("libpluto_amd.so is not built (run __graft_entry__.build() or "
                        "python -m pluto_amd.build)")
    # PyTorch ships its own HIP runtime with the same soname as ROCm's
    # (libamdhip64.so.7) but asks for it by another name.  Loaded first,
    # torch's copy also satisfies this library; loaded after this library,
    # torch would map a second HIP runtime into the process and fail to
    # initialise its devices.  So bring torch in first when it is installed.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    i, d = ctypes.c_int, ctypes.c_double
    sig = {
        "pgpu_last_error": ([], ctypes.c_char_p),
        "pgpu_device_count": ([ctypes.POINTER(i)], i),
        "pgpu_domain_create": ([i, i, i, i, d, d, d, d, ctypes.POINTER(_vp)], i),
        "pgpu_domain_create_bounded": ([i, i, i, i, d, d, d, d, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(_vp)], i),
        "pgpu_domain_destroy": ([_vp], i),
        "pgpu_set_precision": ([_vp, i], i),
        "pgpu_set_pipeline": ([_vp, i], i),
        "pgpu_set_emf_tiling": ([i, i, i], i),
        "pgpu_set_emf_kernel": ([i, i], i),
        "pgpu_set_flux_kernel": ([i, i], i),
        "pgpu_set_solver": ([_vp, i], i),
        "pgpu_set_emf_average": ([_vp, i], i),
        "pgpu_write_dbl_host": ([ctypes.c_char_p, i, i, i, i, i, _dp, _dp, _dp, _dp, d, d,
                                 ctypes.c_long], i),
        "pgpu_read_dbl_host": ([ctypes.c_char_p, i, i, i, i, i, _dp, _dp, _dp, _dp,
                                ctypes.POINTER(d), ctypes.POINTER(d),
                                ctypes.POINTER(ctypes.c_long)], i),
        "pgpu_write_dbl": ([_vp, ctypes.c_char_p, i], i),
        "pgpu_read_dbl": ([_vp, ctypes.c_char_p, i], i),
        "pgpu_get_emf_tiling": ([ctypes.POINTER(i)] * 3, i),
        "pgpu_set_state": ([_vp, _dp, _dp, _dp, _dp], i),
        "pgpu_set_state_bounded": ([_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp], i),
        "pgpu_set_userdef_boundary": ([_vp, USERDEF_FN, _vp], i),
        "pgpu_set_branch_signature": ([_vp, i], i),
        "pgpu_get_branch_signature": ([_vp, ctypes.POINTER(ctypes.c_ulonglong)], i),
        "pgpu_set_branch_tape": ([_vp, i, i], i),
        "pgpu_get_state_bounded": ([_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp], i),
        "pgpu_get_state": ([_vp, _dp, _dp, _dp, _dp], i),
        "pgpu_state_device_ptrs": ([_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)], i),
        "pgpu_set_dt": ([_vp, d, d, d, i], i),
        "pgpu_run": ([_vp, i], i),
        "pgpu_sync": ([_vp], i),
        "pgpu_get_scalars": ([_vp, _dp], i),
        "pgpu_get_history": ([_vp, i, _dp, _dp, _dp], i),
        "pgpu_get_flags": ([_vp, ctypes.POINTER(ctypes.c_ubyte)], i),
        "pgpu_timing_enable": ([_vp, i], i),
        "pgpu_timing_read": ([_vp, i, _dp, ctypes.POINTER(ctypes.c_long)], i),
        "pgpu_kernel_name": ([i, ctypes.POINTER(ctypes.c_char_p)], i),
        "pgpu_time_steps": ([_vp, i, _dp], i),
        "pgpu_domain_bytes": ([_vp, ctypes.POINTER(ctypes.c_size_t)], i),
        "pgpu_state_error": ([_vp, _vp, _dp], i),
        "pgpu_set_state_device": ([_vp, _vp, _vp], i),
        "pgpu_get_state_device": ([_vp, _vp, _vp], i),
        "pgpu_adjoint_step": ([_vp, _vp, _vp], i),
        "pgpu_adjoint_record": ([_vp, _vp], i),
        "pgpu_adjoint_step_taped": ([_vp, _vp, _vp, _vp], i),
        "pgpu_hlld_tangent_rows": ([i, i, d, _dp, _dp, _dp, _dp, _dp, _dp], i),
        "pgpu_hlld_adjoint_rows": ([i, i, d, _dp, _dp, _dp, _dp, _dp], i),
        "pgpu_batch_create": ([i, ctypes.POINTER(i), d, ctypes.POINTER(_vp)], i),
        "pgpu_batch_destroy": ([_vp], i),
        "pgpu_batch_set_patch": ([_vp, i, _dp, _dp, _dp, _dp], i),
        "pgpu_batch_get_patch": ([_vp, i, _dp, _dp, _dp, _dp], i),
        "pgpu_batch_set_dt": ([_vp, d, d, d, i], i),
        "pgpu_batch_run": ([_vp, i], i),
        "pgpu_batch_sync": ([_vp], i),
        "pgpu_batch_time_steps": ([_vp, i, _dp], i),
        "pgpu_batch_patch_scalars": ([_vp, i, _dp], i),
        "pgpu_batch_set_launch": ([_vp, i, i], i),
        "pgpu_batch_get_launch": ([_vp, ctypes.POINTER(i), ctypes.POINTER(i)], i),
        "pgpu_batch_set_tile": ([_vp, i, i, i], i),
        "pgpu_batch_get_tile": ([_vp] + [ctypes.POINTER(i)] * 3, i),
        "pgpu_batch_set_mode": ([_vp, i], i),
        "pgpu_batch_get_mode": ([_vp] + [ctypes.POINTER(i)] * 2, i),
        "pgpu_batch_tune": ([_vp, i, ctypes.POINTER(i), ctypes.POINTER(i), _dp], i),
        "pgpu_batch_bytes": ([_vp, ctypes.POINTER(ctypes.c_size_t)], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise PGPUError("pluto_amd error %d: %s" % (rc, lib().pgpu_last_error().decode()))


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().pgpu_device_count(ctypes.byref(n)))
    return n.value


def write_dbl_host(directory: str, nfile: int, st: MHDState, time: float, dt: float,
                   nstep: int) -> None:
    """Write st as output #nfile in PLUTO's .dbl format (data.NNNN.dbl, the
    dbl.out line, the restart.out record) into directory, on the host."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (st.prim, st.bx, st.by, st.bz)]
    nz, ny, nx = st.shape
    check(lib().pgpu_write_dbl_host(os.fsencode(directory), nfile, nx, ny, nz, st.dims,
                                    *[_p(a) for a in arrs], time, dt, nstep))


def read_dbl_host(directory: str, nfile: int, shape, dims: int = 3, length=None):
    """Read output #nfile of a PLUTO .dbl run: (MHDState, time, dt, nstep)."""
    nz, ny, nx = shape
    prim = np.empty((8, nz, ny, nx))
    bx, by, bz = (np.empty((nz, ny, nx)) for _ in range(3))
    t, dt, ns = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_long(0)
    check(lib().pgpu_read_dbl_host(os.fsencode(directory), nfile, nx, ny, nz, dims, _p(prim),
                                   _p(bx), _p(by), _p(bz), ctypes.byref(t), ctypes.byref(dt),
                                   ctypes.byref(ns)))
    st = MHDState(prim, bx, by, bz, dims, tuple(length) if length else (1.0, 1.0, 1.0))
    return st, t.value, dt.value, ns.value


def set_emf_tiling(bx: int = 0, by: int = 0, kz: int = 0) -> None:
    """Tiling of the z-marching EMF kernel (process-wide); no arguments
    restores the default.  Results do not depend on it."""
    check(lib().pgpu_set_emf_tiling(bx, by, kz))


def set_emf_kernel(kind: int = 3, kz: int = 0) -> None:
    """UCT_HLL EMF kernel of FP64 3D boxes (process-wide): 0 register z-march,
    LDS-staged with tiles 1: 64 x 4, 2: 32 x 4, 3: 32 x 8 (default); kz
    planes per workgroup (0: 16).  Results do not depend on it."""
    check(lib().pgpu_set_emf_kernel(int(kind), int(kz)))


def set_flux_kernel(kind: int = 1, seg: int = 0) -> None:
    """x/y flux kernel of the fused step (process-wide): 0 overlapping
    z-plane tiles, 1 row workgroups marching up y segments of seg rows
    (0: by size; the default).  FP64 results do not depend on it; the
    FP32-flux path's differ in the last FP32 bits (contraction)."""
    check(lib().pgpu_set_flux_kernel(int(kind), int(seg)))


def emf_tiling():
    v = [ctypes.c_int(0) for _ in range(3)]
    check(lib().pgpu_get_emf_tiling(*[ctypes.byref(x) for x in v]))
    return tuple(x.value for x in v)


def _p(a):
    return a.ctypes.data_as(_dp)


class Domain:
    """One periodic box resident on the GPU (see include/pluto_amd.h)."""

    def __init__(self, shape, dims=3, spacing=(1.0, 1.0, 1.0), gamma=5.0 / 3.0,
                 precision=PGPU_PREC_FP64, pipeline=PIPELINE_FUSED, solver=SOLVER_HLLD,
                 emf_average=EMF_UCT_HLL, bounds=None, nghost=2):
        """bounds: None for a periodic box, or the boundary types of the six
        sides (X1_BEG, X1_END, ... X3_END) as BOUND_* codes or PLUTO's
        pluto.ini words ("outflow", "reflective", "axisymmetric",
        "eqtsymmetric", "periodic"): a 3D box held with nghost ghost zones."""
        nz, ny, nx = shape
        self.shape = (nz, ny, nx)
        self.dims = dims
        self.spacing = tuple(float(x) for x in spacing)
        self.gamma = gamma
        h = _vp()
        if bounds is None:
            check(lib().pgpu_domain_create(nx, ny, nz, dims, self.spacing[0], self.spacing[1],
                                           self.spacing[2], gamma, ctypes.byref(h)))
        else:
            codes = [BOUNDS[b] if isinstance(b, str) else int(b) for b in bounds]
            if len(codes) != 6 or dims != 3:
                raise PGPUError("a bounded box is 3D with six boundary types")
            arr = (ctypes.c_int * 6)(*codes)
            check(lib().pgpu_domain_create_bounded(nx, ny, nz, nghost, self.spacing[0],
                                                   self.spacing[1], self.spacing[2], gamma, arr,
                                                   ctypes.byref(h)))
        self.h = h
        if precision != PGPU_PREC_FP64:
            check(lib().pgpu_set_precision(self.h, precision))
        if pipeline != PIPELINE_FUSED:
            check(lib().pgpu_set_pipeline(self.h, pipeline))
        if solver != SOLVER_HLLD:
            self.set_solver(solver)
        if emf_average != EMF_UCT_HLL:
            self.set_emf_average(emf_average)

    def set_emf_average(self, emf_average):
        """EMF average of the CT update (EMF_*, or a key of EMF_AVERAGES)."""
        if isinstance(emf_average, str):
            emf_average = EMF_AVERAGES[emf_average]
        check(lib().pgpu_set_emf_average(self.h, emf_average))

    def set_solver(self, solver):
        """Riemann solver of the sweeps (SOLVER_HLLD, _HLL, _HLLC, _TVDLF or _ROE)."""
        check(lib().pgpu_set_solver(self.h, solver))

    @classmethod
    def from_state(cls, st: MHDState, **kw):
        dom = cls(st.shape, st.dims, st.spacing(), **kw)
        dom.set_state(st)
        return dom

    def close(self):
        if getattr(self, "h", None):
            lib().pgpu_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, st: MHDState, low=None):
        """Upload a state.  low: for a bounded box, the faces below the first
        cells (bx_low[nz][ny], by_low[nz][nx], bz_low[ny][nx]); None starts
        them as images of the last faces."""
        assert st.shape == self.shape
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (st.prim, st.bx, st.by, st.bz)]
        if low is None:
            check(lib().pgpu_set_state(self.h, *[_p(a) for a in arrs]))
            return
        nz, ny, nx = self.shape
        lows = [np.ascontiguousarray(a, dtype=np.float64) for a in low]
        for a, shp in zip(lows, ((nz, ny), (nz, nx), (ny, nx))):
            if a.shape != shp:
                raise PGPUError("lower faces must have shapes (nz,ny), (nz,nx), (ny,nx)")
        check(lib().pgpu_set_state_bounded(self.h, *[_p(a) for a in arrs + lows]))

    def get_state(self, template: MHDState | None = None) -> MHDState:
        nz, ny, nx = self.shape
        prim = np.empty((8, nz, ny, nx))
        bx, by, bz = (np.empty((nz, ny, nx)) for _ in range(3))
        check(lib().pgpu_get_state(self.h, _p(prim), _p(bx), _p(by), _p(bz)))
        length = template.length if template is not None else tuple(
            s * n for s, n in zip(self.spacing, (nx, ny, nz)))
        return MHDState(prim, bx, by, bz, self.dims, tuple(length))

    def set_userdef_boundary(self, fn):
        """Host hook of the USERDEF sides: fn(side, stage, slab) with the
        side 0..5 (X1_BEG .. X3_END), the stage 1 or 2 and a Slab; it
        writes the side's ghost zones (and the transverse cell-centred
        field there) in place.  None removes it."""
        if fn is None:
            self._userdef = None
            check(lib().pgpu_set_userdef_boundary(self.h, USERDEF_FN(), None))
            return

        def tramp(ctx, side, stage, slab):
            try:
                fn(side, stage, slab.contents)
                return 0
            except Exception:        # reported as the hook's failure
                import traceback
                traceback.print_exc()
                return 1

        self._userdef = USERDEF_FN(tramp)    # kept alive with the domain
        check(lib().pgpu_set_userdef_boundary(self.h, self._userdef, None))

    def set_branch_signature(self, on=True):
        """Record (and clear) the branch signature of the steps that follow."""
        check(lib().pgpu_set_branch_signature(self.h, int(bool(on))))

    def branch_signature(self):
        """64-bit hash of the branch outcomes of the recorded steps."""
        h = ctypes.c_ulonglong(0)
        check(lib().pgpu_get_branch_signature(self.h, ctypes.byref(h)))
        return h.value

    def set_branch_tape(self, mode, max_steps=0):
        """Branch tape: TAPE_RECORD (capacity max_steps steps), TAPE_REPLAY
        (the recorded outcomes imposed from the first step on) or TAPE_OFF."""
        check(lib().pgpu_set_branch_tape(self.h, int(mode), int(max_steps)))

    def low_faces(self):
        """A bounded box's faces below the first cells: (bx_low[nz][ny],
        by_low[nz][nx], bz_low[ny][nx])."""
        nz, ny, nx = self.shape
        prim = np.empty((8, nz, ny, nx))
        bx, by, bz = (np.empty((nz, ny, nx)) for _ in range(3))
        low = (np.empty((nz, ny)), np.empty((nz, nx)), np.empty((ny, nx)))
        check(lib().pgpu_get_state_bounded(self.h, _p(prim), _p(bx), _p(by), _p(bz),
                                           *[_p(a) for a in low]))
        return low

    def write_dbl(self, directory: str, nfile: int):
        """Output #nfile of this domain (state, time, next dt, step count)
        in PLUTO's .dbl format."""
        check(lib().pgpu_write_dbl(self.h, os.fsencode(directory), nfile))

    def read_dbl(self, directory: str, nfile: int):
        """Restart from output #nfile: state, time, dt and step count (the
        CFL settings of the last set_dt stay)."""
        check(lib().pgpu_read_dbl(self.h, os.fsencode(directory), nfile))

    def set_dt(self, dt, cfl=0.4, cfl_max_var=1.1, fixed=False):
        check(lib().pgpu_set_dt(self.h, dt, cfl, cfl_max_var, int(bool(fixed))))

    def run(self, nsteps):
        check(lib().pgpu_run(self.h, int(nsteps)))

    def sync(self):
        check(lib().pgpu_sync(self.h))

    def scalars(self):
        out = np.zeros(6)
        check(lib().pgpu_get_scalars(self.h, _p(out)))
        return dict(dt=out[0], invdt=out[1], maxmach=out[2], time=out[3], step=int(out[4]),
                    fail_cells=int(out[5]))

    def history(self, n):
        dt, inv, mach = np.zeros(n), np.zeros(n), np.zeros(n)
        check(lib().pgpu_get_history(self.h, n, _p(dt), _p(inv), _p(mach)))
        return dt, inv, mach

    def flags(self):
        f = np.zeros(self.shape, dtype=np.uint8)
        check(lib().pgpu_get_flags(self.h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))))
        return f

    def timing(self, enable=True):
        check(lib().pgpu_timing_enable(self.h, int(bool(enable))))

    def timing_read(self):
        n = len(KERNEL_IDS)
        ms = np.zeros(n)
        cnt = (ctypes.c_long * n)()
        check(lib().pgpu_timing_read(self.h, n, _p(ms), cnt))
        return {KERNEL_IDS[q]: (float(ms[q]), int(cnt[q])) for q in range(n)}

    def time_steps(self, nsteps):
        ms = ctypes.c_double(0.0)
        check(lib().pgpu_time_steps(self.h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def error_vs(self, ref: "Domain"):
        """Per-field relative L1 and Linf difference of this domain's state
        against ref's: {field: (L1_rel, Linf_rel)}."""
        out = np.zeros(32)
        check(lib().pgpu_state_error(ref.h, self.h, _p(out)))
        res = {}
        for f, name in enumerate(FIELDS):
            s_d, s_r, m_d, m_r = out[4 * f:4 * f + 4]
            res[name] = (s_d / s_r if s_r > 0 else s_d, m_d / m_r if m_r > 0 else m_d)
        return res

    def nbytes(self):
        b = ctypes.c_size_t(0)
        check(lib().pgpu_domain_bytes(self.h, ctypes.byref(b)))
        return b.value


class Batch:
    """Independent periodic patches advanced together (see pgpu_batch_* in
    include/pluto_amd.h).  Patch p is a unit cube of edges[p]^3 cells."""

    def __init__(self, edges, gamma=5.0 / 3.0):
        self.edges = [int(e) for e in edges]
        arr = (ctypes.c_int * len(self.edges))(*self.edges)
        h = _vp()
        check(lib().pgpu_batch_create(len(self.edges), arr, gamma, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().pgpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_patch(self, p, st: MHDState):
        n = self.edges[p]
        assert st.shape == (n, n, n)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (st.prim, st.bx, st.by, st.bz)]
        check(lib().pgpu_batch_set_patch(self.h, p, *[_p(a) for a in arrs]))

    def get_patch(self, p) -> MHDState:
        n = self.edges[p]
        prim = np.empty((8, n, n, n))
        bx, by, bz = (np.empty((n, n, n)) for _ in range(3))
        check(lib().pgpu_batch_get_patch(self.h, p, _p(prim), _p(bx), _p(by), _p(bz)))
        return MHDState(prim, bx, by, bz, 3, (1.0, 1.0, 1.0))

    def set_dt(self, dt, cfl=0.4, cfl_max_var=1.1, fixed=False):
        check(lib().pgpu_batch_set_dt(self.h, dt, cfl, cfl_max_var, int(bool(fixed))))

    def run(self, nsteps):
        check(lib().pgpu_batch_run(self.h, int(nsteps)))

    def sync(self):
        check(lib().pgpu_batch_sync(self.h))

    def time_steps(self, nsteps):
        ms = ctypes.c_double(0.0)
        check(lib().pgpu_batch_time_steps(self.h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def scalars(self, p):
        out = np.zeros(6)
        check(lib().pgpu_batch_patch_scalars(self.h, p, _p(out)))
        return dict(dt=out[0], invdt=out[1], maxmach=out[2], time=out[3], step=int(out[4]),
                    fail_cells=int(out[5]))

    def set_launch(self, threads, groups=0):
        check(lib().pgpu_batch_set_launch(self.h, int(threads), int(groups)))

    def set_mode(self, mode):
        """1: stacks of equal even edges through the fused kernels (default);
        0: every patch through the persistent patch kernel."""
        check(lib().pgpu_batch_set_mode(self.h, int(mode)))

    def mode(self):
        """(mode, number of patches the captured step stacks)"""
        m, n = ctypes.c_int(0), ctypes.c_int(0)
        check(lib().pgpu_batch_get_mode(self.h, ctypes.byref(m), ctypes.byref(n)))
        return m.value, n.value

    def launch(self):
        t, g = ctypes.c_int(0), ctypes.c_int(0)
        check(lib().pgpu_batch_get_launch(self.h, ctypes.byref(t), ctypes.byref(g)))
        return t.value, g.value

    def set_tile(self, tx=0, ty=0, tz=0):
        """Cell tile of a pass (tx*ty*tz = workgroup size), or linear order."""
        check(lib().pgpu_batch_set_tile(self.h, int(tx), int(ty), int(tz)))

    def tile(self):
        t = [ctypes.c_int(0) for _ in range(3)]
        check(lib().pgpu_batch_get_tile(self.h, *[ctypes.byref(x) for x in t]))
        return tuple(x.value for x in t)

    def tune(self, trial_steps=2):
        t, g, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        check(lib().pgpu_batch_tune(self.h, int(trial_steps), ctypes.byref(t), ctypes.byref(g),
                                    ctypes.byref(ms)))
        return t.value, g.value, ms.value

    def nbytes(self):
        b = ctypes.c_size_t(0)
        check(lib().pgpu_batch_bytes(self.h, ctypes.byref(b)))
        return b.value
