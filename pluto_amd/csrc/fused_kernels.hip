// Fused flux kernels of the stage (the fast path used by pgpu_run and
// AdvanceStep).  Same arithmetic as the modular kernels in stage_kernels.hip
// (and thus as the reference), organised to move fewer bytes:
//
//   k_flux_xy<STAGE>  one z-plane tile per workgroup.  The primitive state of
//                     the tile plus its stencil halo is staged in LDS once;
//                     each thread solves one x and one y interface (PLM +
//                     HLLD) from LDS, fluxes go to LDS, and each cell adds
//                     -dt/dx (F_{i+1/2}-F_{i-1/2}) and the y counterpart to
//                     its conservative state in registers.  Writes U, the
//                     UCT_HLL face speeds and velocity slopes of its own
//                     faces/cells, and C_dt's x+y part.
//   k_flux_z<STAGE>   one thread per (i,j) column marching up a z segment:
//                     each cell's z PLM and each z interface is computed
//                     once, the previous interface's flux is carried in
//                     registers; adds the z flux difference to U, finishes
//                     C_dt ((0 + cx) + cy) + cz (update_stage.c:229-235) and
//                     its maximum.
//
// The flux differences are added in x, y, z order as in update_stage.c:134-242.

#include <hip/hip_runtime.h>
#include "mhd_device.hpp"
#include "stage_cells.hpp"
#include "stage_kernels.hpp"
#include "fused_kernels.hpp"

namespace pgpu {

namespace {

// face speed / velocity slope of the EMF stencil at cell c: FP64 values as
// double, FP32 values as float in the first half of the buffer (exact, and
// half the bytes for the EMF kernel to read back)
template <class R>
__device__ __forceinline__ void face_store(double* a, long c, R v)
{
  if constexpr (sizeof(R) == 4) reinterpret_cast<float*>(a)[c] = v;
  else a[c] = v;
}

// what the flux kernels store per face for the EMF: the UCT_HLL data (face
// speeds and velocity slopes), or the face EMFs from the normal fluxes of B
// and the mass flux (ARITHMETIC, UCT0, UCT_CONTACT)
constexpr int FACE_UCT_HLL = 0;
constexpr int FACE_BFLUX = 1;

// periodic index for -n <= q < 2n
__device__ __forceinline__ int wrap(int q, int n)
{
  q = q < 0 ? q + n : q;
  return q >= n ? q - n : q;
}

// periodic index for any q (tiles may be wider than the box)
__device__ __forceinline__ int wrap_any(int q, int n)
{
  q %= n;
  return q < 0 ? q + n : q;
}

__device__ __forceinline__ void block_max(double v, unsigned long long* dst, double* red)
{
  const int lane = threadIdx.x + threadIdx.y * blockDim.x;
  const int wave = lane / warpSize;
  const int nwave = (blockDim.x * blockDim.y + warpSize - 1) / warpSize;
  for (int off = warpSize / 2; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, warpSize);
    v = o > v ? o : v;
  }
  if ((lane % warpSize) == 0) red[wave] = v;
  __syncthreads();
  if (lane == 0) {
    double m = red[0];
    for (int w = 1; w < nwave; w++) m = red[w] > m ? red[w] : m;
    atomicMax(dst, (unsigned long long)__double_as_longlong(m));
  }
}

// PLM of one cell from its neighbours: right-face (vp) and left-face (vm)
// states of all variables.
template <typename R>
__device__ __forceinline__ void plm_cell(const R* vm1, const R* v0, const R* vp1, R* vp, R* vm)
{
#pragma unroll
  for (int q = 0; q < NV; q++) {
    const R s = mc_slope(vm1[q], v0[q], vp1[q]);
    vp[q] = v0[q] + s * R(0.5);
    vm[q] = v0[q] - s * R(0.5);
  }
}

}  // namespace

// ---------------------------------------------------------------------
// x and y sweeps of one z-plane tile
// ---------------------------------------------------------------------

// A TXx8 workgroup produces the (TX-1)x7 cells (i0..i0+TX-2, j0..j0+6) of an
// overlapping tile: thread (tx, ty) solves the x interface on the left of
// cell (i0+tx, j0+ty) and the y interface below cell (i0+tx, j0+ty), so
// every thread does exactly one x and one y interface (two full passes).
// The staged state spans x in [i0-2, i0+TX] and y in [j0-2, j0+8].  TX is
// 32 for FP64 (LDS-limited: three workgroups per CU) and 64 for FP32.
template <int STAGE, typename R, int TX, int S, int E>
__global__ void __launch_bounds__(64 * XY_TY)
k_flux_xy(Geo g, StageIn in, Cons U, FaceData fd, double* cdt_xy, DevScalars* sc)
{
  constexpr int TY = XY_TY;
  constexpr int OX = TX - 1, OY = TY - 1;     // output cells of the workgroup
  constexpr int LX = TX + 3, LY = TY + 3;     // staged region
  constexpr int NT = TX * TY;
  __shared__ R sv[NV][LY][LX];
  __shared__ R fx[6][TY][TX];                 // rho, mx, my, mz, E, p at x_{i0+tx-1/2}
  __shared__ R cmx[TY][TX];
  __shared__ R fy[6][TY][TX];                 // same at y_{j0+ty-1/2}
  __shared__ R cmy[TY][TX];
  __shared__ double red[NT / 64 + 1];

  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = tx + ty * TX;
  const dim3 tb = xcd_tile();
  const int i0 = tb.x * OX, j0 = tb.y * OY, k = tb.z;
  const long plane = (long)k * g.ny * g.nx;

  // stage the primitive state (periodic wrap)
  for (int q = tid; q < LX * LY; q += NT) {
    const int lx = q % LX, ly = q / LX;
    const int gi = wrap_any(i0 + lx - 2, g.nx), gj = wrap_any(j0 + ly - 2, g.ny);
    const long c = plane + (long)gj * g.nx + gi;
#pragma unroll
    for (int v = 0; v < NV; v++) sv[v][ly][lx] = R(in.V[v][c]);
  }
  __syncthreads();

  const double dt = sc->dt;
  const R gam = R(g.gamma), gm1 = R(g.gmm1);
  double mach = 0.0;
  const int ci = i0 + tx, cj = j0 + ty;                 // unwrapped cell of this thread
  const bool row_in = cj < g.ny, col_in = ci < g.nx;

  // x interface between cells (ci-1, cj) and (ci, cj): cells ci-2 .. ci+1 at LDS x = tx .. tx+3
  {
    R a[4][NV], vL[NV], vR[NV], vm0[NV], vpn[NV], vmn[NV], dvc[3];
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int v = 0; v < NV; v++) a[s][v] = sv[v][ty + 2][tx + s];
    plm_cell(a[0], a[1], a[2], vL, vm0);
    plm_cell(a[1], a[2], a[3], vpn, vmn);
#pragma unroll
    for (int v = 0; v < NV; v++) vR[v] = vmn[v];
#pragma unroll
    for (int v = 0; v < 3; v++) dvc[v] = vL[VX1 + v] - vm0[VX1 + v];
    const int oi = wrap_any(ci - 1, g.nx), oj = wrap_any(cj, g.ny);   // owner cell of the face
    const long cf = plane + (long)oj * g.nx + oi;
    const R bface = R(in.b[0][cf]);
    vL[BX1] = bface;
    vR[BX1] = bface;
    RiemannOut<R> r;
    riemann<S, 0, E == FACE_BFLUX>(vL, vR, gam, gm1, r, g.dims == 2);
    fx[0][ty][tx] = r.f[RHO];
    fx[1][ty][tx] = r.f[VX1];
    fx[2][ty][tx] = r.f[VX2];
    fx[3][ty][tx] = r.f[VX3];
    fx[4][ty][tx] = r.f[ENG];
    fx[5][ty][tx] = r.press;
    cmx[ty][tx] = r.cmax;
    // the face belongs to cell ci-1, an output cell of this tile when tx >= 1
    if (tx >= 1 && ty < OY && row_in && ci - 1 < g.nx) {
      if constexpr (E == FACE_BFLUX) {
        fd.SL[0][cf] = -r.f[BX2];   // ezi, eyi (ct_emf.c:131-133)
        fd.SR[0][cf] = r.f[BX3];
        fd.dv[0][0][cf] = r.f[RHO];
      } else {
        face_store(fd.SL[0], cf, max_ge(R(0), -r.SL));
        face_store(fd.SR[0], cf, max_ge(R(0), r.SR));
        face_store(fd.dv[0][0], cf, dvc[0]);
        face_store(fd.dv[0][1], cf, dvc[1]);
        face_store(fd.dv[0][2], cf, dvc[2]);
      }
    }
    const double rm = double(r.mach);
    if (in_sweep(g, 0, ci - 1, cj, k)) mach = rm > mach ? rm : mach;
  }
  // y interface between cells (ci, cj-1) and (ci, cj): rows cj-2 .. cj+1 at LDS y = ty .. ty+3
  {
    R a[4][NV], vL[NV], vR[NV], vm0[NV], vpn[NV], vmn[NV], dvc[3];
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int v = 0; v < NV; v++) a[s][v] = sv[v][ty + s][tx + 2];
    plm_cell(a[0], a[1], a[2], vL, vm0);
    plm_cell(a[1], a[2], a[3], vpn, vmn);
#pragma unroll
    for (int v = 0; v < NV; v++) vR[v] = vmn[v];
#pragma unroll
    for (int v = 0; v < 3; v++) dvc[v] = vL[VX1 + v] - vm0[VX1 + v];
    const int oi = wrap_any(ci, g.nx), oj = wrap_any(cj - 1, g.ny);   // owner cell of the face
    const long cf = plane + (long)oj * g.nx + oi;
    const R bface = R(in.b[1][cf]);
    vL[BX2] = bface;
    vR[BX2] = bface;
    RiemannOut<R> r;
    riemann<S, 1, E == FACE_BFLUX>(vL, vR, gam, gm1, r, g.dims == 2);
    fy[0][ty][tx] = r.f[RHO];
    fy[1][ty][tx] = r.f[VX1];
    fy[2][ty][tx] = r.f[VX2];
    fy[3][ty][tx] = r.f[VX3];
    fy[4][ty][tx] = r.f[ENG];
    fy[5][ty][tx] = r.press;
    cmy[ty][tx] = r.cmax;
    // the face belongs to cell cj-1, an output cell of this tile when ty >= 1
    if (ty >= 1 && tx < OX && col_in && cj - 1 < g.ny) {
      if constexpr (E == FACE_BFLUX) {
        fd.SL[1][cf] = r.f[BX1];    // ezj, exj (ct_emf.c:153-155)
        fd.SR[1][cf] = -r.f[BX3];
        fd.dv[1][0][cf] = r.f[RHO];
      } else {
        face_store(fd.SL[1], cf, max_ge(R(0), -r.SL));
        face_store(fd.SR[1], cf, max_ge(R(0), r.SR));
        face_store(fd.dv[1][0], cf, dvc[0]);
        face_store(fd.dv[1][1], cf, dvc[1]);
        face_store(fd.dv[1][2], cf, dvc[2]);
      }
    }
    const double rm = double(r.mach);
    if (in_sweep(g, 1, ci, cj - 1, k)) mach = rm > mach ? rm : mach;
  }
  __syncthreads();

  // output cells: base state plus x then y flux differences
  if (tx < OX && ty < OY && col_in && row_in) {
    const long c = plane + (long)cj * g.nx + ci;
    double u[5];
    if (STAGE == 1) {
      double v[NV];
      if constexpr (sizeof(R) == sizeof(double)) {
#pragma unroll
        for (int q = 0; q < NV; q++) v[q] = sv[q][ty + 2][tx + 2];
      } else {
#pragma unroll
        for (int q = 0; q < NV; q++) v[q] = in.V[q][c];   // U0 from the FP64 state
      }
      u[0] = v[RHO];
      u[1] = v[RHO] * v[VX1];
      u[2] = v[RHO] * v[VX2];
      u[3] = v[RHO] * v[VX3];
      u[4] = total_energy(v, g.gmm1);
    } else {
#pragma unroll
      for (int q = 0; q < 5; q++) u[q] = U.u[q][c];
    }
    const double dtdx = dt / g.dx[0], dtdy = dt / g.dx[1];
    double r[5];
#pragma unroll
    for (int q = 0; q < 5; q++) r[q] = -dtdx * (double(fx[q][ty][tx + 1]) - double(fx[q][ty][tx]));
    r[1] -= dtdx * (double(fx[5][ty][tx + 1]) - double(fx[5][ty][tx]));
#pragma unroll
    for (int q = 0; q < 5; q++) u[q] += r[q];
#pragma unroll
    for (int q = 0; q < 5; q++) r[q] = -dtdy * (double(fy[q][ty + 1][tx]) - double(fy[q][ty][tx]));
    r[2] -= dtdy * (double(fy[5][ty + 1][tx]) - double(fy[5][ty][tx]));
#pragma unroll
    for (int q = 0; q < 5; q++) u[q] += r[q];
#pragma unroll
    for (int q = 0; q < 5; q++) U.u[q][c] = u[q];
    if (STAGE == 1) {
      double cdt = 0.0;
      cdt += 0.5 * (double(cmx[ty][tx]) + double(cmx[ty][tx + 1])) * g.inv_dx[0];
      cdt += 0.5 * (double(cmy[ty][tx]) + double(cmy[ty + 1][tx])) * g.inv_dx[1];
      cdt_xy[c] = cdt;
    }
  }
  block_max(mach, &sc->mach_max, red);
}

// Non-overlapping variant (the FP32-flux path): a 32x8 tile plus a two-cell
// halo; the (TX+1)*TY x and TX*(TY+1) y interfaces of the tile are spread
// over the workgroup (three passes).  With FP32 state the workgroup is not
// LDS-limited and this layout measured faster than the overlapping one.
template <int STAGE, typename R, int S>
__global__ void __launch_bounds__(XY_TX * XY_TY)
k_flux_xy_items(Geo g, StageIn in, Cons U, FaceData fd, double* cdt_xy, DevScalars* sc)
{
  constexpr int TX = XY_TX, TY = XY_TY;
  constexpr int LX = TX + 4, LY = TY + 4;     // staged region with 2-cell halo
  constexpr int NXI = (TX + 1) * TY;          // x interfaces of the tile (one halo)
  constexpr int NYI = TX * (TY + 1);          // y interfaces
  constexpr int NT = TX * TY;
  __shared__ R sv[NV][LY][LX];
  __shared__ R fx[6][TY][TX + 1];             // rho, mx, my, mz, E, p
  __shared__ R cmx[TY][TX + 1];
  __shared__ R fy[6][TY + 1][TX];
  __shared__ R cmy[TY + 1][TX];
  __shared__ double red[NT / 64 + 1];

  const int tid = threadIdx.x + threadIdx.y * TX;
  const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY, k = blockIdx.z;
  const long plane = (long)k * g.ny * g.nx;

  // stage the primitive state of the tile plus halo (periodic wrap)
  for (int q = tid; q < LX * LY; q += NT) {
    const int lx = q % LX, ly = q / LX;
    const int gi = wrap_any(i0 + lx - 2, g.nx), gj = wrap_any(j0 + ly - 2, g.ny);
    const long c = plane + (long)gj * g.nx + gi;
#pragma unroll
    for (int v = 0; v < NV; v++) sv[v][ly][lx] = R(in.V[v][c]);
  }
  __syncthreads();

  const double dt = sc->dt;
  const R gam = R(g.gamma), gm1 = R(g.gmm1);
  double mach = 0.0;

  // interfaces: x items first, then y items; items that no cell of the
  // box needs (tiles wider than the box) are skipped
  for (int item = tid; item < NXI + NYI; item += NT) {
    RiemannOut<R> r;
    R vL[NV], vR[NV], dvc[3];
    R a[4][NV];
    int li, lj, own, gi, gj;
    if (item < NXI) {
      const int row = item / (TX + 1), col = item % (TX + 1);
      if (j0 + row >= g.ny || i0 + col - 1 >= g.nx) continue;
      // interface between tile cells col-1 and col of row `row`
      li = col + 1;   // LDS x of the left cell
      lj = row + 2;
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int v = 0; v < NV; v++) a[s][v] = sv[v][lj][li - 1 + s];
      gi = i0 + col - 1;
      gj = j0 + row;
      own = (col >= 1) && (gi < g.nx) && (gj < g.ny);
      R vm0[NV], vpn[NV], vmn[NV];
      plm_cell(a[0], a[1], a[2], vL, vm0);
      plm_cell(a[1], a[2], a[3], vpn, vmn);
#pragma unroll
      for (int v = 0; v < NV; v++) vR[v] = vmn[v];
#pragma unroll
      for (int v = 0; v < 3; v++) dvc[v] = vL[VX1 + v] - vm0[VX1 + v];
      const long cf = plane + (long)wrap(gj, g.ny) * g.nx + wrap(gi, g.nx);
      const R bface = R(in.b[0][cf]);
      vL[BX1] = bface;
      vR[BX1] = bface;
      riemann<S, 0, false>(vL, vR, gam, gm1, r, g.dims == 2);
      fx[0][row][col] = r.f[RHO];
      fx[1][row][col] = r.f[VX1];
      fx[2][row][col] = r.f[VX2];
      fx[3][row][col] = r.f[VX3];
      fx[4][row][col] = r.f[ENG];
      fx[5][row][col] = r.press;
      cmx[row][col] = r.cmax;
      if (own) {
        const long c = plane + (long)gj * g.nx + gi;
        face_store(fd.SL[0], c, max_ge(R(0), -r.SL));
        face_store(fd.SR[0], c, max_ge(R(0), r.SR));
        face_store(fd.dv[0][0], c, dvc[0]);
        face_store(fd.dv[0][1], c, dvc[1]);
        face_store(fd.dv[0][2], c, dvc[2]);
      }
    } else {
      const int it = item - NXI;
      const int row = it / TX, col = it % TX;   // row 0 is the halo interface j0-1/2
      if (i0 + col >= g.nx || j0 + row - 1 >= g.ny) continue;
      li = col + 2;
      lj = row + 1;
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int v = 0; v < NV; v++) a[s][v] = sv[v][lj - 1 + s][li];
      gi = i0 + col;
      gj = j0 + row - 1;
      own = (row >= 1) && (gi < g.nx) && (gj < g.ny);
      R vm0[NV], vpn[NV], vmn[NV];
      plm_cell(a[0], a[1], a[2], vL, vm0);
      plm_cell(a[1], a[2], a[3], vpn, vmn);
#pragma unroll
      for (int v = 0; v < NV; v++) vR[v] = vmn[v];
#pragma unroll
      for (int v = 0; v < 3; v++) dvc[v] = vL[VX1 + v] - vm0[VX1 + v];
      const long cf = plane + (long)wrap(gj, g.ny) * g.nx + wrap(gi, g.nx);
      const R bface = R(in.b[1][cf]);
      vL[BX2] = bface;
      vR[BX2] = bface;
      riemann<S, 1, false>(vL, vR, gam, gm1, r, g.dims == 2);
      fy[0][row][col] = r.f[RHO];
      fy[1][row][col] = r.f[VX1];
      fy[2][row][col] = r.f[VX2];
      fy[3][row][col] = r.f[VX3];
      fy[4][row][col] = r.f[ENG];
      fy[5][row][col] = r.press;
      cmy[row][col] = r.cmax;
      if (own) {
        const long c = plane + (long)gj * g.nx + gi;
        face_store(fd.SL[1], c, max_ge(R(0), -r.SL));
        face_store(fd.SR[1], c, max_ge(R(0), r.SR));
        face_store(fd.dv[1][0], c, dvc[0]);
        face_store(fd.dv[1][1], c, dvc[1]);
        face_store(fd.dv[1][2], c, dvc[2]);
      }
    }
    const double rm = double(r.mach);
    mach = rm > mach ? rm : mach;
  }
  __syncthreads();

  // per cell: base state plus x then y flux differences
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i = i0 + tx, j = j0 + ty;
  if (i < g.nx && j < g.ny) {
    const long c = plane + (long)j * g.nx + i;
    double u[5];
    if (STAGE == 1) {
      double v[NV];
      if constexpr (sizeof(R) == sizeof(double)) {
#pragma unroll
        for (int q = 0; q < NV; q++) v[q] = sv[q][ty + 2][tx + 2];
      } else {
#pragma unroll
        for (int q = 0; q < NV; q++) v[q] = in.V[q][c];   // U0 from the FP64 state
      }
      u[0] = v[RHO];
      u[1] = v[RHO] * v[VX1];
      u[2] = v[RHO] * v[VX2];
      u[3] = v[RHO] * v[VX3];
      u[4] = total_energy(v, g.gmm1);
    } else {
#pragma unroll
      for (int q = 0; q < 5; q++) u[q] = U.u[q][c];
    }
    const double dtdx = dt / g.dx[0], dtdy = dt / g.dx[1];
    double r[5];
#pragma unroll
    for (int q = 0; q < 5; q++) r[q] = -dtdx * (double(fx[q][ty][tx + 1]) - double(fx[q][ty][tx]));
    r[1] -= dtdx * (double(fx[5][ty][tx + 1]) - double(fx[5][ty][tx]));
#pragma unroll
    for (int q = 0; q < 5; q++) u[q] += r[q];
#pragma unroll
    for (int q = 0; q < 5; q++) r[q] = -dtdy * (double(fy[q][ty + 1][tx]) - double(fy[q][ty][tx]));
    r[2] -= dtdy * (double(fy[5][ty + 1][tx]) - double(fy[5][ty][tx]));
#pragma unroll
    for (int q = 0; q < 5; q++) u[q] += r[q];
#pragma unroll
    for (int q = 0; q < 5; q++) U.u[q][c] = u[q];
    if (STAGE == 1) {
      double cdt = 0.0;
      cdt += 0.5 * (double(cmx[ty][tx]) + double(cmx[ty][tx + 1])) * g.inv_dx[0];
      cdt += 0.5 * (double(cmy[ty][tx]) + double(cmy[ty + 1][tx])) * g.inv_dx[1];
      cdt_xy[c] = cdt;
    }
  }
  block_max(mach, &sc->mach_max, red);
}

// ---------------------------------------------------------------------
// z sweep, one thread per column marching over a z segment
// ---------------------------------------------------------------------

template <int STAGE, typename R, int S, int E>
__global__ void __launch_bounds__(Z_TX * Z_TY)
k_flux_z(Geo g, StageIn in, Cons U, FaceData fd, const double* cdt_xy, int seg, DevScalars* sc)
{
  __shared__ double red[(Z_TX * Z_TY) / 64 + 1];
  const int i = blockIdx.x * Z_TX + threadIdx.x;
  const int j = blockIdx.y * Z_TY + threadIdx.y;
  const int k0 = blockIdx.z * seg;
  const int k1 = min(k0 + seg, g.nz);
  double mach = 0.0, cdt_max = 0.0;
  if (i < g.nx && j < g.ny) {
    const long col = (long)j * g.nx + i;
    const long pl = (long)g.nx * g.ny;
    const double dt = sc->dt;
    const double dtdz = dt / g.dx[2];
    const R gam = R(g.gamma), gm1 = R(g.gmm1);
    auto load = [&](int kk, R* v) {
      const long c = (long)wrap(kk, g.nz) * pl + col;
#pragma unroll
      for (int q = 0; q < NV; q++) v[q] = R(in.V[q][c]);
    };
    R va[NV], vb[NV], vc[NV];
    R vp_k[NV], vm_k[NV];
    // PLM of cell k0-1 and the interface (k0-1)+1/2
    load(k0 - 2, va);
    load(k0 - 1, vb);
    load(k0, vc);
    plm_cell(va, vb, vc, vp_k, vm_k);     // cell k0-1
    R fprev[6], cmprev;
    {
      R vnext[NV], vp_n[NV], vm_n[NV];
      load(k0 + 1, vnext);
      plm_cell(vb, vc, vnext, vp_n, vm_n);   // cell k0
      R vL[NV], vR[NV];
#pragma unroll
      for (int q = 0; q < NV; q++) {
        vL[q] = vp_k[q];
        vR[q] = vm_n[q];
      }
      const R bface = R(in.b[2][(long)wrap(k0 - 1, g.nz) * pl + col]);
      vL[BX3] = bface;
      vR[BX3] = bface;
      RiemannOut<R> r;
      riemann<S, 2, false>(vL, vR, gam, gm1, r);
      if (in_sweep(g, 2, i, j, k0 - 1)) mach = double(r.mach);
      fprev[0] = r.f[RHO]; fprev[1] = r.f[VX1]; fprev[2] = r.f[VX2];
      fprev[3] = r.f[VX3]; fprev[4] = r.f[ENG]; fprev[5] = r.press;
      cmprev = r.cmax;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        vp_k[q] = vp_n[q];     // cell k0
        vm_k[q] = vm_n[q];
        va[q] = vb[q];
        vb[q] = vc[q];
        vc[q] = vnext[q];
      }
    }
    // march: va = v(k-1), vb = v(k), vc = v(k+1), vp_k/vm_k = PLM of cell k
    for (int k = k0; k < k1; k++) {
      R vn[NV], vp_n[NV], vm_n[NV];
      load(k + 2, vn);
      plm_cell(vb, vc, vn, vp_n, vm_n);     // cell k+1
      R vL[NV], vR[NV];
#pragma unroll
      for (int q = 0; q < NV; q++) {
        vL[q] = vp_k[q];
        vR[q] = vm_n[q];
      }
      const long c = (long)k * pl + col;
      const R bface = R(in.b[2][c]);
      vL[BX3] = bface;
      vR[BX3] = bface;
      RiemannOut<R> r;
      riemann<S, 2, E == FACE_BFLUX>(vL, vR, gam, gm1, r);
      const double rm = double(r.mach);
      if (in_sweep(g, 2, i, j, k)) mach = rm > mach ? rm : mach;
      const R fcur[6] = {r.f[RHO], r.f[VX1], r.f[VX2], r.f[VX3], r.f[ENG], r.press};
      double u[5];
#pragma unroll
      for (int q = 0; q < 5; q++) u[q] = U.u[q][c];
      double rr[5];
#pragma unroll
      for (int q = 0; q < 5; q++) rr[q] = -dtdz * (double(fcur[q]) - double(fprev[q]));
      rr[3] -= dtdz * (double(fcur[5]) - double(fprev[5]));
#pragma unroll
      for (int q = 0; q < 5; q++) u[q] += rr[q];
#pragma unroll
      for (int q = 0; q < 5; q++) U.u[q][c] = u[q];
      if constexpr (E == FACE_BFLUX) {
        fd.SL[2][c] = -r.f[BX1];    // eyk, exk (ct_emf.c:175-176)
        fd.SR[2][c] = r.f[BX2];
        fd.dv[2][0][c] = r.f[RHO];
      } else {
        face_store(fd.SL[2], c, max_ge(R(0), -r.SL));
        face_store(fd.SR[2], c, max_ge(R(0), r.SR));
        face_store(fd.dv[2][0], c, vp_k[VX1] - vm_k[VX1]);
        face_store(fd.dv[2][1], c, vp_k[VX2] - vm_k[VX2]);
        face_store(fd.dv[2][2], c, vp_k[VX3] - vm_k[VX3]);
      }
      if (STAGE == 1) {
        double cdt = cdt_xy[c];
        cdt += 0.5 * (double(cmprev) + double(r.cmax)) * g.inv_dx[2];
        if (in_dom(g, i, j, k)) cdt_max = cdt > cdt_max ? cdt : cdt_max;
      }
#pragma unroll
      for (int q = 0; q < 6; q++) fprev[q] = fcur[q];
      cmprev = r.cmax;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        vp_k[q] = vp_n[q];
        vm_k[q] = vm_n[q];
        va[q] = vb[q];
        vb[q] = vc[q];
        vc[q] = vn[q];
      }
    }
  }
  block_max(mach, &sc->mach_max, red);
  if (STAGE == 1) {
    __syncthreads();
    block_max(cdt_max, &sc->invdt_max, red);
  }
}

// 2D problems: C_dt's maximum directly from the x+y part.
__global__ void k_cdt_max2d(Geo g, const double* cdt_xy, DevScalars* sc)
{
  __shared__ double red[4];
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = c < g.n ? cdt_xy[c] : 0.0;
  block_max(v, &sc->invdt_max, red);
}

// ---------------------------------------------------------------------
// CT + face average + ConsToPrim
// ---------------------------------------------------------------------

// Updated face along axis A of the cell at offsets (oi, oj, ok) (<= 0).
template <int A, int STAGE>
__device__ __forceinline__ double new_face(const Geo& g, const EdgeOut& e, const FaceIn& bin,
                                           const FaceIn& b0, int i, int j, int k, double dtx,
                                           double dty, double dtz)
{
  const long c = ((long)k * g.ny + j) * g.nx + i;
  const long cim = c + (i == 0 ? g.nx - 1 : -1);
  const long cjm = c + (j == 0 ? (long)(g.ny - 1) * g.nx : -(long)g.nx);
  double r;
  if (g.dims == 3) {
    const long ckm = c + (k == 0 ? (long)(g.nz - 1) * g.nx * g.ny : -(long)g.nx * g.ny);
    if (A == 0) r = 0.0 - dty * (e.ez[c] - e.ez[cjm]) + dtz * (e.ey[c] - e.ey[ckm]);
    else if (A == 1) r = dtx * (e.ez[c] - e.ez[cim]) - dtz * (e.ex[c] - e.ex[ckm]);
    else r = -dtx * (e.ey[c] - e.ey[cim]) + dty * (e.ex[c] - e.ex[cjm]);
  } else {
    if (A == 0) r = 0.0 - dty * (e.ez[c] - e.ez[cjm]);
    else if (A == 1) r = dtx * (e.ez[c] - e.ez[cim]);
    else r = 0.0;
  }
  double b = (A == 2 && g.dims != 3) ? bin.b[A][c] : bin.b[A][c] + r;
  if (STAGE == 2) b = 0.5 * b0.b[A][c] + 0.5 * b;
  return b;
}

template <int STAGE>
__global__ void __launch_bounds__(256) k_ct_c2p(Geo g, EdgeOut e, FaceIn bin, FaceIn b0,
                                                FaceOut bout, Cons U, StageIn vn, Prims Vout,
                                                unsigned char* flag, DevScalars* sc)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const long c = ((long)k * g.ny + j) * g.nx + i;
  const double dt = sc->dt;
  const double dtx = dt / g.dx[0], dty = dt / g.dx[1], dtz = dt / g.dx[2];
  const int im = (i == 0 ? g.nx - 1 : i - 1), jm = (j == 0 ? g.ny - 1 : j - 1);
  const int km = (k == 0 ? g.nz - 1 : k - 1);
  const double bx = new_face<0, STAGE>(g, e, bin, b0, i, j, k, dtx, dty, dtz);
  const double by = new_face<1, STAGE>(g, e, bin, b0, i, j, k, dtx, dty, dtz);
  const double bz = new_face<2, STAGE>(g, e, bin, b0, i, j, k, dtx, dty, dtz);
  const double bxm = new_face<0, STAGE>(g, e, bin, b0, im, j, k, dtx, dty, dtz);
  const double bym = new_face<1, STAGE>(g, e, bin, b0, i, jm, k, dtx, dty, dtz);
  bout.b[0][c] = bx;
  bout.b[1][c] = by;
  bout.b[2][c] = bz;
  double u[NV], v[NV];
  u[RHO] = U.u[0][c];
  u[VX1] = U.u[1][c];
  u[VX2] = U.u[2][c];
  u[VX3] = U.u[3][c];
  u[ENG] = U.u[4][c];
  if (STAGE == 2) {
    double v0[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) v0[q] = vn.V[q][c];
    const double e0 = total_energy(v0, g.gmm1);
    u[RHO] = 0.5 * v0[RHO] + 0.5 * u[RHO];
    u[VX1] = 0.5 * (v0[RHO] * v0[VX1]) + 0.5 * u[VX1];
    u[VX2] = 0.5 * (v0[RHO] * v0[VX2]) + 0.5 * u[VX2];
    u[VX3] = 0.5 * (v0[RHO] * v0[VX3]) + 0.5 * u[VX3];
    u[ENG] = 0.5 * e0 + 0.5 * u[ENG];
  }
  u[BX1] = 0.5 * (bx + bxm);
  u[BX2] = 0.5 * (by + bym);
  if (g.dims == 3) {
    const double bzm = new_face<2, STAGE>(g, e, bin, b0, i, j, km, dtx, dty, dtz);
    u[BX3] = 0.5 * (bz + bzm);
  } else {
    u[BX3] = 0.0;
  }
  const unsigned char f = cons2prim(u, v, g.gmm1);
#pragma unroll
  for (int q = 0; q < NV; q++) Vout.v[q][c] = v[q];
  if (STAGE == 1) {
    U.u[0][c] = u[RHO];
    U.u[1][c] = u[VX1];
    U.u[2][c] = u[VX2];
    U.u[3][c] = u[VX3];
    U.u[4][c] = u[ENG];
  }
  if (f) {
    flag[c] |= f;
    if (in_dom(g, i, j, k)) atomicAdd(&sc->fail_cells, 1u);
  }
  if (has_nan(v) && in_dom(g, i, j, k)) atomicAdd(STAGE == 1 ? &sc->nan_cells : &sc->nan_next, 1u);
}

// z-marching CT + ConsToPrim (3D, fields under 4 GiB): a thread owns the
// column (i, j) over kz planes.  The face bz(i, j, k-1) the cell average
// needs is the column's own bz of the previous plane, and the k-1 edge
// values ey(0,0,-1), ey(-1,0,-1), ex(0,0,-1), ex(0,-1,-1) are the previous
// plane's ey(0,0,0), ey(-1,0,0), ex(0,0,0), ex(0,-1,0): both stay in
// registers, so a cell reads 31 values in stage 2 (18 in stage 1) rather than
// 37 (23).  Same arithmetic as k_ct_c2p (ct_update.c:87-196,
// ct_field_average.c:64-66, rk_step.c:168-174, mappers.c:88-254).
template <int STAGE>
__device__ __forceinline__ double ct_face(double bin, double b0, double r)
{
  double b = bin + r;
  if (STAGE == 2) b = 0.5 * b0 + 0.5 * b;
  return b;
}

template <int STAGE>
__global__ void __launch_bounds__(256) k_ct_c2p_z(Geo g, EdgeOut e, FaceIn bin, FaceIn b0,
                                                  FaceOut bout, Cons U, StageIn vn, Prims Vout,
                                                  unsigned char* flag, DevScalars* sc, int kz)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const double dt = sc->dt;
  const double dtx = dt / g.dx[0], dty = dt / g.dx[1], dtz = dt / g.dx[2];
  const unsigned rowb = (unsigned)g.nx * 8u;
  const unsigned planeb = rowb * (unsigned)g.ny;
  const unsigned o00 = (unsigned)j * rowb + (unsigned)i * 8u;
  const unsigned om0 = (unsigned)j * rowb + (unsigned)wrap2(i, -1, g.nx) * 8u;   // (i-1, j)
  const unsigned o0m = (unsigned)wrap2(j, -1, g.ny) * rowb + (unsigned)i * 8u;   // (i, j-1)
  const unsigned omm = (unsigned)wrap2(j, -1, g.ny) * rowb + (unsigned)wrap2(i, -1, g.nx) * 8u;
  const int k0 = blockIdx.z * kz;
  const int k1 = min(k0 + kz, g.nz);
  auto B0 = [&](int a, unsigned o) { return STAGE == 2 ? at_byte(b0.b[a], o) : 0.0; };

  // plane k0-1: edge values and the face bz
  double ey_p, eym_p, ex_p, exm_p, bz_p;
  {
    const unsigned pk = (unsigned)wrap2(k0, -1, g.nz) * planeb;
    ey_p = at_byte(e.ey, pk + o00);
    eym_p = at_byte(e.ey, pk + om0);
    ex_p = at_byte(e.ex, pk + o00);
    exm_p = at_byte(e.ex, pk + o0m);
    const double rz = -dtx * (ey_p - eym_p) + dty * (ex_p - exm_p);
    bz_p = ct_face<STAGE>(at_byte(bin.b[2], pk + o00), B0(2, pk + o00), rz);
  }
  for (int k = k0; k < k1; k++) {
    const unsigned pk = (unsigned)k * planeb;
    const unsigned c = pk + o00;
    const double ez = at_byte(e.ez, c), ez_jm = at_byte(e.ez, pk + o0m);
    const double ez_im = at_byte(e.ez, pk + om0), ez_imjm = at_byte(e.ez, pk + omm);
    const double ey = at_byte(e.ey, c), ey_im = at_byte(e.ey, pk + om0);
    const double ex = at_byte(e.ex, c), ex_jm = at_byte(e.ex, pk + o0m);
    // own faces (i+1/2, j+1/2, k+1/2) and the low faces (i-1/2), (j-1/2)
    const double bx = ct_face<STAGE>(at_byte(bin.b[0], c), B0(0, c),
                                      0.0 - dty * (ez - ez_jm) + dtz * (ey - ey_p));
    const double by = ct_face<STAGE>(at_byte(bin.b[1], c), B0(1, c),
                                      dtx * (ez - ez_im) - dtz * (ex - ex_p));
    const double bz = ct_face<STAGE>(at_byte(bin.b[2], c), B0(2, c),
                                      -dtx * (ey - ey_im) + dty * (ex - ex_jm));
    const double bxm = ct_face<STAGE>(at_byte(bin.b[0], pk + om0), B0(0, pk + om0),
                                       0.0 - dty * (ez_im - ez_imjm) + dtz * (ey_im - eym_p));
    const double bym = ct_face<STAGE>(at_byte(bin.b[1], pk + o0m), B0(1, pk + o0m),
                                       dtx * (ez_jm - ez_imjm) - dtz * (ex_jm - exm_p));
    at_byte(bout.b[0], c) = bx;
    at_byte(bout.b[1], c) = by;
    at_byte(bout.b[2], c) = bz;
    double u[NV], v[NV];
    u[RHO] = at_byte(U.u[0], c);
    u[VX1] = at_byte(U.u[1], c);
    u[VX2] = at_byte(U.u[2], c);
    u[VX3] = at_byte(U.u[3], c);
    u[ENG] = at_byte(U.u[4], c);
    if (STAGE == 2) {
      double v0[NV];
#pragma unroll
      for (int q = 0; q < NV; q++) v0[q] = at_byte(vn.V[q], c);
      const double e0 = total_energy(v0, g.gmm1);
      u[RHO] = 0.5 * v0[RHO] + 0.5 * u[RHO];
      u[VX1] = 0.5 * (v0[RHO] * v0[VX1]) + 0.5 * u[VX1];
      u[VX2] = 0.5 * (v0[RHO] * v0[VX2]) + 0.5 * u[VX2];
      u[VX3] = 0.5 * (v0[RHO] * v0[VX3]) + 0.5 * u[VX3];
      u[ENG] = 0.5 * e0 + 0.5 * u[ENG];
    }
    u[BX1] = 0.5 * (bx + bxm);
    u[BX2] = 0.5 * (by + bym);
    u[BX3] = 0.5 * (bz + bz_p);
    const unsigned char f = cons2prim(u, v, g.gmm1);
#pragma unroll
    for (int q = 0; q < NV; q++) at_byte(Vout.v[q], c) = v[q];
    if (STAGE == 1) {
      at_byte(U.u[0], c) = u[RHO];
      at_byte(U.u[1], c) = u[VX1];
      at_byte(U.u[2], c) = u[VX2];
      at_byte(U.u[3], c) = u[VX3];
      at_byte(U.u[4], c) = u[ENG];
    }
    if (f) {
      flag[(unsigned long)(c >> 3)] |= f;
      if (in_dom(g, i, j, k)) atomicAdd(&sc->fail_cells, 1u);
    }
    if (has_nan(v) && in_dom(g, i, j, k))
      atomicAdd(STAGE == 1 ? &sc->nan_cells : &sc->nan_next, 1u);
    ey_p = ey;
    eym_p = ey_im;
    ex_p = ex;
    exm_p = ex_jm;
    bz_p = bz;
  }
}

int launch_ct_c2p(const Geo& g, int stage, const EdgeOut& e, const FaceIn& bin, const FaceIn& b0,
                  const FaceOut& bout, const Cons& U, const StageIn& vn, const Prims& Vout,
                  unsigned char* flag, DevScalars* sc, hipStream_t s, KernelTimer* kt)
{
  const dim3 b(64, 4), gr((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
  if (kt) kt->begin(KT_CT, s);
  if (g.dims == 3 && (unsigned long long)g.n * 8ull < (1ull << 32)) {
    // planes per workgroup: up to 16, fewer while the box would get fewer
    // than 2048 workgroups
    const long xy = (long)gr.x * gr.y;
    int kz = 16;
    while (kz > 1 && xy * ((g.nz + kz - 1) / kz) < 2048) kz /= 2;
    const dim3 gz(gr.x, gr.y, (g.nz + kz - 1) / kz);
    if (stage == 1)
      hipLaunchKernelGGL(k_ct_c2p_z<1>, gz, b, 0, s, g, e, bin, b0, bout, U, vn, Vout, flag, sc, kz);
    else
      hipLaunchKernelGGL(k_ct_c2p_z<2>, gz, b, 0, s, g, e, bin, b0, bout, U, vn, Vout, flag, sc, kz);
  } else if (stage == 1) {
    hipLaunchKernelGGL(k_ct_c2p<1>, gr, b, 0, s, g, e, bin, b0, bout, U, vn, Vout, flag, sc);
  } else {
    hipLaunchKernelGGL(k_ct_c2p<2>, gr, b, 0, s, g, e, bin, b0, bout, U, vn, Vout, flag, sc);
  }
  if (kt) kt->end(KT_CT, s);
  return (int)hipGetLastError();
}

template <typename R, int S, int E = FACE_UCT_HLL>
static void launch_flux_fused_t(const Geo& g, int stage, const StageIn& in, const Cons& U,
                                const FaceData& fd, double* cdt_xy, DevScalars* sc,
                                hipStream_t s, KernelTimer* kt)
{
  if (kt) kt->begin(KT_FLUX_XY, s);
  if constexpr (sizeof(R) == sizeof(double)) {
    constexpr int TX = 32;
    const dim3 bxy(TX, XY_TY), gxy((g.nx + TX - 2) / (TX - 1), (g.ny + XY_TY - 2) / (XY_TY - 1), g.nz);
    if (stage == 1) hipLaunchKernelGGL((k_flux_xy<1, R, TX, S, E>), gxy, bxy, 0, s, g, in, U, fd, cdt_xy, sc);
    else hipLaunchKernelGGL((k_flux_xy<2, R, TX, S, E>), gxy, bxy, 0, s, g, in, U, fd, cdt_xy, sc);
  } else {
    static_assert(E == FACE_UCT_HLL, "the FP32 x/y kernel stores UCT_HLL face data only");
    const dim3 bxy(XY_TX, XY_TY), gxy((g.nx + XY_TX - 1) / XY_TX, (g.ny + XY_TY - 1) / XY_TY, g.nz);
    if (stage == 1) hipLaunchKernelGGL((k_flux_xy_items<1, R, S>), gxy, bxy, 0, s, g, in, U, fd, cdt_xy, sc);
    else hipLaunchKernelGGL((k_flux_xy_items<2, R, S>), gxy, bxy, 0, s, g, in, U, fd, cdt_xy, sc);
  }
  if (kt) kt->end(KT_FLUX_XY, s);
  if (g.dims == 3) {
    // z segment per column: 32 planes (one extra interface per segment),
    // halved down to 16 while the box would get fewer than 512 workgroups
    // (measured: 256^3 best at 32, 128^3 at 16)
    const long xy = (long)((g.nx + Z_TX - 1) / Z_TX) * ((g.ny + Z_TY - 1) / Z_TY);
    int seg = g.nz >= 64 ? 32 : (g.nz >= 16 ? 8 : g.nz);
    if (seg == 32 && xy * ((g.nz + seg - 1) / seg) < 512) seg = 16;
    const dim3 bz(Z_TX, Z_TY), gz((g.nx + Z_TX - 1) / Z_TX, (g.ny + Z_TY - 1) / Z_TY,
                                  (g.nz + seg - 1) / seg);
    if (kt) kt->begin(KT_FLUX_Z, s);
    if (stage == 1)
      hipLaunchKernelGGL((k_flux_z<1, R, S, E>), gz, bz, 0, s, g, in, U, fd, cdt_xy, seg, sc);
    else
      hipLaunchKernelGGL((k_flux_z<2, R, S, E>), gz, bz, 0, s, g, in, U, fd, cdt_xy, seg, sc);
    if (kt) kt->end(KT_FLUX_Z, s);
  } else if (stage == 1) {
    hipLaunchKernelGGL(k_cdt_max2d, dim3((unsigned)((g.n + 255) / 256)), dim3(256), 0, s, g,
                       cdt_xy, sc);
  }
}

int launch_flux_fused(const Geo& g, int stage, bool fp32_flux, const StageIn& in, const Cons& U,
                      const FaceData& fd, double* cdt_xy, DevScalars* sc, hipStream_t s,
                      KernelTimer* kt)
{
  if (g.emf_avg != EMF_UCT_HLL) {
    // face EMFs for the ARITHMETIC / UCT0 / UCT_CONTACT averages (FP64 only)
    if (fp32_flux) return (int)hipErrorInvalidValue;
    if (g.solver == SOLVER_HLL)
      launch_flux_fused_t<double, SOLVER_HLL, FACE_BFLUX>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
    else if (g.solver == SOLVER_HLLC)
      launch_flux_fused_t<double, SOLVER_HLLC, FACE_BFLUX>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
    else if (g.solver == SOLVER_TVDLF)
      launch_flux_fused_t<double, SOLVER_TVDLF, FACE_BFLUX>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
    else if (g.solver == SOLVER_ROE)
      launch_flux_fused_t<double, SOLVER_ROE, FACE_BFLUX>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
    else
      launch_flux_fused_t<double, SOLVER_HLLD, FACE_BFLUX>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
    return (int)hipGetLastError();
  }
  if (g.solver == SOLVER_HLL) {
    if (fp32_flux) launch_flux_fused_t<float, SOLVER_HLL>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
    else launch_flux_fused_t<double, SOLVER_HLL>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  } else if (g.solver == SOLVER_HLLC) {
    if (fp32_flux) launch_flux_fused_t<float, SOLVER_HLLC>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
    else launch_flux_fused_t<double, SOLVER_HLLC>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  } else if (g.solver == SOLVER_TVDLF) {
    if (fp32_flux) launch_flux_fused_t<float, SOLVER_TVDLF>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
    else launch_flux_fused_t<double, SOLVER_TVDLF>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  } else if (g.solver == SOLVER_ROE) {
    if (fp32_flux) launch_flux_fused_t<float, SOLVER_ROE>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
    else launch_flux_fused_t<double, SOLVER_ROE>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  } else {
    if (fp32_flux) launch_flux_fused_t<float, SOLVER_HLLD>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
    else launch_flux_fused_t<double, SOLVER_HLLD>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  }
  return (int)hipGetLastError();
}

}  // namespace pgpu
