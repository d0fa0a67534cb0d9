"""Initial conditions for the unsplit MHD stage (host side, numpy).

A state is held as a :class:`MHDState`: cell-centred primitives in the
canonical order ``rho, vx, vy, vz, Bx, By, Bz, p`` (PLUTO's NVAR order for
3D MHD, ``MHD/mod_defs.h:20-41``) and the three staggered face fields
``bx[k,j,i]`` (at ``x_{i+1/2}``), ``by`` (at ``y_{j+1/2}``), ``bz`` (at
``z_{k+1/2}``) on a periodic box without ghost zones.

Magnetic fields are built the way PLUTO builds them from a vector potential:
``A`` is sampled on cell edges and the face fields are its discrete curl
(``vec_pot_diff.c:146-169``), so the discrete divergence is zero to
round-off.  The cell-centred field is the face average
``0.5*(b_{i+1/2} + b_{i-1/2})`` (``startup.c:235-241``).
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

GAMMA = 5.0 / 3.0

RHO, VX, VY, VZ, BX, BY, BZ, PRS = range(8)
NAMES = ("rho", "vx", "vy", "vz", "Bx", "By", "Bz", "p")


@dataclasses.dataclass
class MHDState:
    prim: np.ndarray          # (8, nz, ny, nx) float64
    bx: np.ndarray            # (nz, ny, nx)
    by: np.ndarray
    bz: np.ndarray
    dims: int = 3             # 2 for an x-y problem (nz == 1, no z sweep)
    length: tuple = (1.0, 1.0, 1.0)   # box size (Lx, Ly, Lz)

    @property
    def shape(self):
        return self.prim.shape[1:]

    @property
    def nx(self):
        return self.prim.shape[3]

    @property
    def ny(self):
        return self.prim.shape[2]

    @property
    def nz(self):
        return self.prim.shape[1]

    def copy(self) -> "MHDState":
        return MHDState(self.prim.copy(), self.bx.copy(), self.by.copy(),
                        self.bz.copy(), self.dims, tuple(self.length))

    def spacing(self):
        """Cell widths as PLUTO computes them: ``(xR-xL)/N`` (set_grid.c:399)."""
        nz, ny, nx = self.shape
        lx, ly, lz = self.length
        return (lx / nx, ly / ny, lz / nz)


def face_average(bx, by, bz, dims=3):
    """Cell-centred field from faces, ``0.5*(b_i + b_{i-1})`` per axis."""
    cbx = 0.5 * (bx + np.roll(bx, 1, axis=2))
    cby = 0.5 * (by + np.roll(by, 1, axis=1))
    if dims == 3:
        cbz = 0.5 * (bz + np.roll(bz, 1, axis=0))
    else:
        cbz = np.zeros_like(bz)
    return cbx, cby, cbz


def curl_edges(ax, ay, az, d, dims=3):
    """Face fields from edge-centred vector potential, PLUTO's discrete curl.

    ``ax[k,j,i]`` sits at (x_i, y_{j+1/2}, z_{k+1/2}), ``ay`` at
    (x_{i+1/2}, y_j, z_{k+1/2}), ``az`` at (x_{i+1/2}, y_{j+1/2}, z_k).
    Follows vec_pot_diff.c:139-169 term for term (same operation order).
    """
    dx, dy, dz = d
    bx = (az - np.roll(az, 1, axis=1)) / dy
    by = -(az - np.roll(az, 1, axis=2)) / dx
    if dims == 3:
        bx = bx - (ay - np.roll(ay, 1, axis=0)) / dz
        by = by + (ax - np.roll(ax, 1, axis=0)) / dz
        bz = (ay - np.roll(ay, 1, axis=2)) / dx - (ax - np.roll(ax, 1, axis=1)) / dy
    else:
        bz = np.zeros_like(az)
    return bx, by, bz


def divb(st: MHDState) -> np.ndarray:
    """Discrete divergence of the face field per cell (ct_update.c:281-313)."""
    dx, dy, dz = st.spacing()
    d = (st.bx - np.roll(st.bx, 1, axis=2)) / dx + (st.by - np.roll(st.by, 1, axis=1)) / dy
    if st.dims == 3:
        d = d + (st.bz - np.roll(st.bz, 1, axis=0)) / dz
    return d


def assemble(rho, v, p, bx, by, bz, dims, length) -> MHDState:
    nz, ny, nx = rho.shape
    prim = np.empty((8, nz, ny, nx))
    prim[RHO] = rho
    prim[VX:VZ + 1] = v
    prim[PRS] = p
    cbx, cby, cbz = face_average(bx, by, bz, dims)
    prim[BX], prim[BY], prim[BZ] = cbx, cby, cbz
    return MHDState(prim, np.ascontiguousarray(bx), np.ascontiguousarray(by),
                    np.ascontiguousarray(bz), dims, tuple(length))


# --------------------------------------------------------------------------
# 2D Orszag-Tang vortex on [0,1]^2
# --------------------------------------------------------------------------

def orszag_tang_2d(n: int) -> MHDState:
    """Orszag-Tang vortex, periodic [0,1]^2, gamma = 5/3.

    rho = 25/(36 pi), p = 5/(12 pi), v = (-sin 2pi y, sin 2pi x),
    A_z = (cos(2 pi y)/(2 pi) + cos(4 pi x)/(4 pi)) / sqrt(4 pi) on z-edges,
    giving B = (-sin 2pi y, sin 4pi x)/sqrt(4 pi) with div B = 0 discretely.
    This is the unit-box form of Test_Problems/MHD/Orszag_Tang/init.c.
    """
    dx = dy = 1.0 / n
    x = (np.arange(n) + 0.5) * dx          # cell centres
    xr = (np.arange(n) + 1.0) * dx         # right faces
    X, Y = np.meshgrid(x, x, indexing="xy")
    XR, YR = np.meshgrid(xr, xr, indexing="xy")
    s = 1.0 / math.sqrt(4.0 * math.pi)
    rho = np.full((1, n, n), 25.0 / (36.0 * math.pi))
    p = np.full((1, n, n), 5.0 / (12.0 * math.pi))
    v = np.zeros((3, 1, n, n))
    v[0, 0] = -np.sin(2.0 * math.pi * Y)
    v[1, 0] = np.sin(2.0 * math.pi * X)
    az = s * (np.cos(2.0 * math.pi * YR) / (2.0 * math.pi)
              + np.cos(4.0 * math.pi * XR) / (4.0 * math.pi))
    az = az[None]
    zero = np.zeros_like(az)
    bx, by, bz = curl_edges(zero, zero, az, (dx, dy, 1.0), dims=2)
    return assemble(rho, v, p, bx, by, bz, 2, (1.0, 1.0, 1.0))


# --------------------------------------------------------------------------
# Seeded Gaussian random fields (3D)
# --------------------------------------------------------------------------

def _mode_table(kmin, kmax, slope):
    r = int(math.floor(kmax))
    ks = np.arange(-r, r + 1)
    KZ, KY, KX = np.meshgrid(ks, ks, ks, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    sel = (kk >= kmin) & (kk <= kmax)
    kx, ky, kz, kabs = KX[sel], KY[sel], KZ[sel], kk[sel]
    amp = kabs ** (-slope)
    return kx, ky, kz, amp


def random_field(n, rng, kmax=8.0, kmin=1.0, slope=2.0, shift=(0.0, 0.0, 0.0)):
    """Real Gaussian random field on an n^3 periodic grid of the unit box.

    f(x) = Re sum_k a_k c_k exp(2 pi i k.x) over integer wavevectors with
    kmin <= |k| <= kmax, amplitude a_k = |k|^-slope (amplitude spectrum), c_k
    complex standard normal drawn from ``rng`` in a fixed mode order.  The
    field is evaluated at cell centres displaced by ``shift`` (in cells), so
    edge- and face-centred samples of one field are exact band-limited
    evaluations.  Not normalised.
    """
    kx, ky, kz, amp = _mode_table(kmin, kmax, slope)
    c = rng.standard_normal(kx.size) + 1j * rng.standard_normal(kx.size)
    return _synth(n, kx, ky, kz, amp * c, shift)


def _synth(n, kx, ky, kz, coef, shift):
    spec = np.zeros((n, n, n), dtype=np.complex128)
    # positions x = (i + 0.5 + shift)/n
    ph = np.exp(2j * math.pi * (kx * (0.5 + shift[0]) + ky * (0.5 + shift[1])
                                + kz * (0.5 + shift[2])) / n)
    np.add.at(spec, (kz % n, ky % n, kx % n), coef * ph)
    try:
        import scipy.fft as sfft
        f = sfft.ifftn(spec, workers=-1)
    except Exception:  # pragma: no cover - scipy is present in this image
        f = np.fft.ifftn(spec)
    return np.ascontiguousarray((f.real * (n ** 3)))


class _Field:
    """One random field's coefficients; sample() evaluates it anywhere."""

    def __init__(self, rng, kmax, kmin=1.0, slope=2.0):
        self.kx, self.ky, self.kz, amp = _mode_table(kmin, kmax, slope)
        c = rng.standard_normal(self.kx.size) + 1j * rng.standard_normal(self.kx.size)
        self.coef = amp * c

    def sample(self, n, shift=(0.0, 0.0, 0.0)):
        return _synth(n, self.kx, self.ky, self.kz, self.coef, shift)


def _unit(f):
    f = f - f.mean()
    return f / f.std()


def random_mhd_3d(n: int, seed: int = 2024, kmax: float = 8.0, slope: float = 2.0,
                  drho: float = 0.1, dp: float = 0.05, v_rms: float = 0.5,
                  b_rms: float = 0.5) -> MHDState:
    """Random-field MHD state on the periodic unit cube, gamma = 5/3.

    rho = 1 + drho*g, p = 1 + dp*g' with g, g' unit-variance Gaussian random
    fields; v (3 components) with rms speed ``v_rms``; A on cell edges,
    scaled so the discrete curl has rms |B| = ``b_rms``.  All fields use
    integer wavevectors 1 <= |k| <= kmax with amplitude |k|^-slope and a
    numpy PCG64 stream seeded with ``seed`` (one draw sequence: g, g', vx,
    vy, vz, Ax, Ay, Az).
    """
    rng = np.random.default_rng(seed)
    d = (1.0 / n,) * 3
    fields = [_Field(rng, kmax, 1.0, slope) for _ in range(8)]
    g = _unit(fields[0].sample(n))
    gp = _unit(fields[1].sample(n))
    v = np.stack([fields[2 + c].sample(n) for c in range(3)])
    v -= v.mean(axis=(1, 2, 3), keepdims=True)
    v *= v_rms / math.sqrt(np.mean(np.sum(v * v, axis=0)))
    ax = fields[5].sample(n, (0.0, 0.5, 0.5))
    ay = fields[6].sample(n, (0.5, 0.0, 0.5))
    az = fields[7].sample(n, (0.5, 0.5, 0.0))
    bx, by, bz = curl_edges(ax, ay, az, d, 3)
    cbx, cby, cbz = face_average(bx, by, bz, 3)
    scale = b_rms / math.sqrt(np.mean(cbx ** 2 + cby ** 2 + cbz ** 2))
    bx, by, bz = curl_edges(ax * scale, ay * scale, az * scale, d, 3)
    rho = 1.0 + drho * g
    p = 1.0 + dp * gp
    return assemble(rho, v, p, bx, by, bz, 3, (1.0, 1.0, 1.0))


def smooth_mhd_controls(n: int, seed: int = 5, kmax: float = 4.0, amp: float = 0.05,
                        slope: float = 2.0):
    """Control fields of the smooth state: rho, v (3, n, n, n) and the
    edge-centred vector-potential perturbation (ax, ay, az), each a seeded
    |k| <= kmax random field normalised to rms ``amp`` (rho about 1)."""
    rng = np.random.default_rng(seed)
    fields = [_Field(rng, kmax, 1.0, slope) for _ in range(7)]

    def norm(f):
        f = f - f.mean()
        return f * (amp / math.sqrt(np.mean(f * f)))

    rho = 1.0 + norm(fields[0].sample(n))
    v = np.stack([norm(fields[1 + c].sample(n)) for c in range(3)])
    # normalise A by the rms of its own samples at the edge positions
    ax = fields[4].sample(n, (0.0, 0.5, 0.5))
    ay = fields[5].sample(n, (0.5, 0.0, 0.5))
    az = fields[6].sample(n, (0.5, 0.5, 0.0))
    ax, ay, az = (a * (amp / math.sqrt(np.mean(a * a))) for a in (ax, ay, az))
    return rho, v, (ax, ay, az)


def smooth_mhd_3d(n: int, seed: int = 5, kmax: float = 4.0, amp: float = 0.05,
                  slope: float = 2.0) -> MHDState:
    """Uniform rho=1, p=1, B=(1,0,0) plus seeded |k|<=kmax perturbations.

    rho, each velocity component and each vector-potential component get a
    random field normalised to rms ``amp``; B = (1,0,0) + curl(dA).
    """
    d = (1.0 / n,) * 3
    rho, v, (ax, ay, az) = smooth_mhd_controls(n, seed, kmax, amp, slope)
    bx, by, bz = curl_edges(ax, ay, az, d, 3)
    bx = bx + 1.0
    p = np.ones_like(rho)
    return assemble(rho, v, p, bx, by, bz, 3, (1.0, 1.0, 1.0))


def uniform_state(shape, rho=1.0, v=(0.0, 0.0, 0.0), b=(0.0, 0.0, 0.0), p=1.0,
                  dims=3, length=(1.0, 1.0, 1.0)) -> MHDState:
    nz, ny, nx = shape
    prim = np.empty((8, nz, ny, nx))
    prim[RHO] = rho
    prim[VX], prim[VY], prim[VZ] = v
    prim[BX], prim[BY], prim[BZ] = b
    prim[PRS] = p
    bx = np.full(shape, float(b[0]))
    by = np.full(shape, float(b[1]))
    bz = np.full(shape, float(b[2]))
    return MHDState(prim, bx, by, bz, dims, tuple(length))


def blast_3d(n: int, p_in: float = 100.0, p_out: float = 0.1, radius: float = 0.1,
             b0: float = 1.0) -> MHDState:
    """Magnetised blast: rho = 1, p = p_in inside r < radius of the box centre,
    p_out outside, uniform B = b0 (1,1,1)/sqrt(3) (face values exact, so
    div B = 0).  Strong shocks exercise every HLLD region."""
    x = (np.arange(n) + 0.5) / n - 0.5
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    rho = np.ones((n, n, n))
    p = np.where(r < radius, p_in, p_out)
    v = np.zeros((3, n, n, n))
    b = b0 / math.sqrt(3.0)
    bx = np.full((n, n, n), b)
    by = np.full((n, n, n), b)
    bz = np.full((n, n, n), b)
    return assemble(rho, v, p, bx, by, bz, 3, (1.0, 1.0, 1.0))
