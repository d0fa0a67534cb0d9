"""ctypes mirrors of PLUTO's structs and array layouts, for calling the
library's PLUTO entry points from Python (the tests, the bench) the way a
PLUTO build calls them.

Layouts follow the reference: ARRAY_4D / ARRAY_3D pointer tables over one
contiguous block (arrays.c), ArrayBox with a lower index of -1 for the
staggered fields (arrays.c:341, initialize.c:448-453), Uc with the variable
index fastest, Sweep rows with stateR offset by one (tools.c:161-231).
"""
import ctypes

import numpy as np

vp = ctypes.c_void_p


class State(ctypes.Structure):
    _fields_ = [(n, vp) for n in ("v", "u", "flux", "fluxCR", "lambda_", "Bbck", "prs", "a2",
                                  "cw", "h", "J", "cCR", "Fcr", "Lp", "Rp")] + [
        ("fill", ctypes.c_char * 8)]


class Sweep(ctypes.Structure):
    _fields_ = [(n, vp) for n in ("vn", "flux", "tc_flux", "lmax", "src", "rhs", "press", "bn",
                                  "SL", "SR", "flag")] + [
        ("stateL", State), ("stateR", State), ("stateC", State), ("fill", ctypes.c_char * 40)]


class Grid(ctypes.Structure):
    _fields_ = [("xbeg", ctypes.c_double * 3), ("xend", ctypes.c_double * 3),
                ("xbeg_glob", ctypes.c_double * 3), ("xend_glob", ctypes.c_double * 3)] + [
        (n, vp * 3) for n in ("x", "x_glob", "xr", "xr_glob", "xl", "xl_glob", "dx", "dx_glob",
                              "xgc")] + [
        ("dV", vp), ("A", vp * 3), ("dx_dl", vp * 3), ("rt", vp), ("sp", vp), ("s", vp),
        ("dmu", vp), ("inv_dx", vp * 3), ("inv_dxi", vp * 3), ("dl_min", ctypes.c_double * 3)] + [
        (n, ctypes.c_int * 3) for n in ("np_tot_glob", "np_int_glob", "np_tot", "np_int", "nghost",
                                        "lbound", "rbound", "gbeg", "gend", "beg", "end", "lbeg",
                                        "lend", "uniform", "nproc", "rank_coord")] + [
        ("level", ctypes.c_int), ("fill", ctypes.c_char * 376)]


class TimeStep(ctypes.Structure):
    _fields_ = [("cmax", vp)] + [(n, ctypes.c_double) for n in (
        "invDt_hyp", "invDt_par", "invDt_particles", "dt_cool", "cfl", "cfl_par", "rmax_par",
        "clock_particles", "clock_particles_bound", "clock_hyp", "clock_par", "clock_cooling",
        "clock_tot")] + [(n, ctypes.c_int) for n in ("Nsub_particles", "Nsts", "Nrkc", "Nrkl")]


class Data(ctypes.Structure):
    _fields_ = [(n, vp) for n in ("Vc", "Uc", "Vs", "Vuser", "Ax1", "Ax2", "Ax3", "J", "Tc",
                                  "flag", "PHead", "Fcr", "Jcr", "qcr", "Ecr", "Vdust", "Fdust",
                                  "pstr", "Ex1", "Ex2", "Ex3", "emf", "Dts", "Ft")] + [
        ("fill", ctypes.c_char * 78)]


EMF_ARRAYS = ("exj", "exk", "eyi", "eyk", "ezi", "ezj")
EMF_ARRAYS2 = ("SxL", "SxR", "SyL", "SyR", "SzL", "SzR", "ex", "ey", "ez", "dbx_dy", "dbx_dz",
               "dby_dx", "dby_dz", "dbz_dx", "dbz_dy", "dvx_dx", "dvy_dx", "dvz_dx", "dvx_dy",
               "dvy_dy", "dvz_dy", "dvx_dz", "dvy_dz", "dvz_dz")


class EMF(ctypes.Structure):
    _fields_ = [(n, vp) for n in EMF_ARRAYS + ("svx", "svy", "svz")] + [
        (n, ctypes.c_int) for n in ("ibeg", "jbeg", "kbeg", "iend", "jend", "kend")] + [
        (n, vp) for n in EMF_ARRAYS2]


assert ctypes.sizeof(EMF) == 288 and EMF.ibeg.offset == 72 and EMF.dvx_dx.offset == 216
assert ctypes.sizeof(State) == 128 and ctypes.sizeof(Sweep) == 512
assert ctypes.sizeof(Grid) == 1048 and ctypes.sizeof(TimeStep) == 128
assert ctypes.sizeof(Data) == 272


def _addr(a, idx=0):
    return a.ctypes.data + idx * a.itemsize


class Keep:
    """Holds every numpy buffer a pointer table refers to."""

    def __init__(self):
        self.bufs = []

    def __call__(self, a):
        self.bufs.append(a)
        return a


def rows_table(keep, block, offset_rows=0):
    """double** over the rows of a 2D block; entry -offset_rows is row 0."""
    n = block.shape[0]
    t = keep(np.array([_addr(block, i * block.shape[1]) for i in range(n)], dtype=np.uint64))
    return _addr(t, offset_rows)


def table3(keep, block, lo=(0, 0, 0)):
    """T*** over block[k][j][i] with lower indices lo (each 0 or -1)."""
    nk, nj, ni = block.shape
    isz = block.itemsize
    r = (np.int64(block.ctypes.data) + np.arange(nk * nj, dtype=np.int64) * (ni * isz)
         - lo[2] * isz)
    rows = keep(r.astype(np.uint64).reshape(nk, nj))
    p = np.int64(rows.ctypes.data) + np.arange(nk, dtype=np.int64) * (nj * 8) - lo[1] * 8
    planes = keep(p.astype(np.uint64))
    return _addr(planes) - 8 * lo[0]


class PlutoBox:
    """PLUTO-layout arrays for an interior box with nghost = 2 ghost zones."""

    def __init__(self, st, ng=2):
        self.keep = Keep()
        k = self.keep
        nz, ny, nx = st.shape
        self.st = st
        self.ng = ng
        self.tot = (nx + 2 * ng, ny + 2 * ng, nz + 2 * ng)
        T1, T2, T3 = self.tot
        # Vc: ARRAY_4D(NVAR, NX3_TOT, NX2_TOT, NX1_TOT)
        self.vc = k(np.full((8, T3, T2, T1), np.nan))
        top = k(np.array([table3(k, self.vc[q]) for q in range(8)], dtype=np.uint64))
        self.Vc = _addr(top)
        # Uc: ARRAY_4D(NX3_TOT, NX2_TOT, NX1_TOT, NVAR)
        self.uc = k(np.full((T3, T2, T1, 8), np.nan))
        cells = k((np.int64(self.uc.ctypes.data) +
                   np.arange(T3 * T2 * T1, dtype=np.int64) * 64).astype(np.uint64)
                  .reshape(T3, T2, T1))
        self.Uc = table3(k, cells)
        # Vs: ArrayBox with -1 in the staggered direction
        self.vs = [k(np.full((T3, T2, T1 + 1), np.nan)), k(np.full((T3, T2 + 1, T1), np.nan)),
                   k(np.full((T3 + 1, T2, T1), np.nan))]
        vs_t = [table3(k, self.vs[0], (0, 0, -1)), table3(k, self.vs[1], (0, -1, 0)),
                table3(k, self.vs[2], (-1, 0, 0))]
        self.vs_top = k(np.array(vs_t, dtype=np.uint64))
        self.Vs = _addr(self.vs_top)
        self.fl = k(np.zeros((T3, T2, T1), dtype=np.uint8))
        self.data = Data()
        self.data.Vc, self.data.Uc, self.data.Vs = self.Vc, self.Uc, self.Vs
        self.data.flag = table3(k, self.fl)
        # grid
        g = Grid()
        dxs = st.spacing()
        for a, (n, d) in enumerate(zip((nx, ny, nz), dxs)):
            tot = n + 2 * ng
            dx = k(np.full(tot, d))
            g.dx[a] = dx.ctypes.data
            g.np_int[a] = n
            g.np_tot[a] = tot
            g.nghost[a] = ng
            g.lbeg[a] = g.beg[a] = ng
            g.lend[a] = g.end[a] = ng + n - 1
            g.lbound[a] = g.rbound[a] = 5     # PERIODIC
        self.grid = g
        self.dts = TimeStep()
        self.cmax = k(np.zeros(max(self.tot) + 8))
        self.dts.cmax = self.cmax.ctypes.data
        self.set_state(st)

    def interior(self, a):
        g = self.ng
        return a[..., g:-g, g:-g, g:-g]

    def set_state(self, st):
        g = self.ng
        nz, ny, nx = st.shape
        self.vc[:, g:g + nz, g:g + ny, g:g + nx] = st.prim
        self.vs[0][g:g + nz, g:g + ny, g + 1:g + 1 + nx] = st.bx
        self.vs[1][g:g + nz, g + 1:g + 1 + ny, g:g + nx] = st.by
        self.vs[2][g + 1:g + 1 + nz, g:g + ny, g:g + nx] = st.bz
        # left boundary faces as periodic images (they feed the face average)
        self.vs[0][g:g + nz, g:g + ny, g] = st.bx[:, :, -1]
        self.vs[1][g:g + nz, g, g:g + nx] = st.by[:, -1, :]
        self.vs[2][g, g:g + ny, g:g + nx] = st.bz[-1, :, :]

    def get_state(self):
        from pluto_amd.problems import MHDState
        g = self.ng
        nz, ny, nx = self.st.shape
        prim = self.vc[:, g:g + nz, g:g + ny, g:g + nx].copy()
        bx = self.vs[0][g:g + nz, g:g + ny, g + 1:g + 1 + nx].copy()
        by = self.vs[1][g:g + nz, g + 1:g + 1 + ny, g:g + nx].copy()
        bz = self.vs[2][g + 1:g + 1 + nz, g:g + ny, g:g + nx].copy()
        return MHDState(prim, bx, by, bz, self.st.dims, self.st.length)

    def fill_ghosts(self):
        """Periodic fill of Vc and Vs, as Boundary() does."""
        g = self.ng
        for ax in (1, 2, 3):
            n = self.vc.shape[ax] - 2 * g
            sl = [slice(None)] * 4
            src = [slice(None)] * 4
            for gg in range(g):
                sl[ax] = gg
                src[ax] = gg + n
                self.vc[tuple(sl)] = self.vc[tuple(src)]
                sl[ax] = g + n + gg
                src[ax] = g + gg
                self.vc[tuple(sl)] = self.vc[tuple(src)]


class SweepRows:
    """A Sweep whose rows are numpy arrays ([n][8])."""

    def __init__(self, nmax):
        self.keep = Keep()
        k = self.keep
        self.nmax = nmax
        self.sw = Sweep()
        mk = lambda: k(np.full((nmax + 2, 8), np.nan))
        self.vC, self.vL, self.vR, self.uL, self.uR, self.flux = (mk() for _ in range(6))
        self.press, self.bn, self.SL, self.SR = (k(np.full(nmax + 2, np.nan)) for _ in range(4))
        self.sw.stateC.v = rows_table(k, self.vC)
        self.sw.stateL.v = rows_table(k, self.vL)
        self.sw.stateL.u = rows_table(k, self.uL)
        # stateR arrays are allocated with +1 so stateR.v[-1] is valid
        self.sw.stateR.v = rows_table(k, self.vR, 1)
        self.sw.stateR.u = rows_table(k, self.uR, 1)
        self.sw.flux = rows_table(k, self.flux)
        self.sw.press = self.press.ctypes.data
        self.sw.bn = self.bn.ctypes.data
        self.sw.SL = self.SL.ctypes.data
        self.sw.SR = self.SR.ctypes.data

    def stateR_v(self, i):
        return self.vR[i + 1]

    def stateR_u(self, i):
        return self.uR[i + 1]


def load_abi(lib_path):
    lib = ctypes.CDLL(lib_path)
    P = ctypes.POINTER
    lib.HLLD_Solver.argtypes = [P(Sweep), ctypes.c_int, ctypes.c_int, vp, vp]
    lib.HLL_Solver.argtypes = [P(Sweep), ctypes.c_int, ctypes.c_int, vp, vp]
    lib.HLLC_Solver.argtypes = [P(Sweep), ctypes.c_int, ctypes.c_int, vp, vp]
    lib.LF_Solver.argtypes = [P(Sweep), ctypes.c_int, ctypes.c_int, vp, vp]
    lib.Roe_Solver.argtypes = [P(Sweep), ctypes.c_int, ctypes.c_int, vp, vp]
    lib.States.argtypes = [P(Sweep), ctypes.c_int, ctypes.c_int, vp]
    lib.PrimToCons.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    lib.ConsToPrim.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp]
    lib.ConsToPrim.restype = ctypes.c_int
    lib.AdvanceStep.argtypes = [P(Data), vp, P(TimeStep), P(Grid)]
    lib.AdvanceStep.restype = ctypes.c_int
    lib.UpdateStage.argtypes = [P(Data), vp, vp, vp, ctypes.c_double, P(TimeStep), P(Grid)]
    lib.CT_AverageMagneticField.argtypes = [vp, vp, P(Grid)]
    lib.SetSolver.argtypes = [ctypes.c_char_p]
    lib.SetSolver.restype = vp
    lib.PGPU_SetEMFAverage.argtypes = [ctypes.c_int]
    lib.PGPU_SetEMFAverage.restype = ctypes.c_int
    lib.pgpu_finalize.argtypes = []
    lib.pgpu_finalize.restype = None
    bind_ct(lib)
    return lib


def bind_ct(lib):
    """argtypes of the CT entry points (ct.h:94-112), in this library or the
    reference's row library."""
    P = ctypes.POINTER
    lib.CT_Allocate.argtypes = [P(EMF)]
    lib.CT_StoreUpwindEMF.argtypes = [P(Sweep), P(EMF), ctypes.c_int, ctypes.c_int, vp]
    lib.CT_StoreVelSlopes.argtypes = [P(EMF), P(Sweep), ctypes.c_int, ctypes.c_int]
    lib.CT_ComputeEMF.argtypes = [P(Data), vp]
    lib.CT_Update.argtypes = [P(Data), vp, ctypes.c_double, vp]
    return lib


def bind_boundary(lib):
    """argtypes of Boundary (boundary.c:40), in this library or the
    reference's row library."""
    lib.Boundary.argtypes = [ctypes.POINTER(Data), ctypes.c_int, vp]
    lib.Boundary.restype = None
    return lib
