// Flux kernels of the fused stage (k_flux_xy_rows, k_flux_xy,
// k_flux_xy_items, k_flux_z) and their launcher template, shared by
// fused_kernels.hip (the FP64 instantiations, compiled without FMA
// contraction: bit-exact with the reference), fused_f32.hip (the FP32-flux
// instantiations, compiled with contraction) and fused_fast.hip (the
// tolerance-mode FP64 instantiations).
// Fused flux kernels of the stage (the fast path used by pgpu_run and
// AdvanceStep).  Same arithmetic as the modular kernels in stage_kernels.hip
// (and thus as the reference), organised to move fewer bytes:
//
//   k_flux_xy_rows    (the default x/y kernel) one workgroup per box row,
//                     marching up a y segment: each cell reconstructed once
//                     per direction, x neighbours through LDS, y carried in
//                     registers (see the kernel).
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

#pragma once
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
// V: 0 for the builds of fused_kernels.hip and fused_f32.hip, 1 for the
// tolerance-mode FP64 build of fused_fast.hip (distinct kernel symbols)
template <int STAGE, typename R, int TX, int S, int E, int V = 0>
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
  sc = scal(sc, g, k);

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
  sc = scal(sc, g, k);

  // stage the primitive state of the tile plus halo (periodic wrap); the
  // second stage's input was stored as float by the first stage
  for (int q = tid; q < LX * LY; q += NT) {
    const int lx = q % LX, ly = q / LX;
    const int gi = wrap_any(i0 + lx - 2, g.nx), gj = wrap_any(j0 + ly - 2, g.ny);
    const long c = plane + (long)gj * g.nx + gi;
#pragma unroll
    for (int v = 0; v < NV; v++)
      sv[v][ly][lx] = (STAGE == 2 && sizeof(R) == 4) ? R(reinterpret_cast<const float*>(in.V[v])[c])
                                                     : R(in.V[v][c]);
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
// x and y sweeps, one workgroup per row marching up a y segment
// ---------------------------------------------------------------------

// The variables a cell exchanges with its x neighbours: all but B_x (the x
// interfaces take the staggered face value instead of the reconstruction).
constexpr int XROW_NV = NV - 1;
__device__ __forceinline__ constexpr int xrow_var(int q) { return q < BX1 ? q : q + 1; }

// The workgroup holds one whole box row (thread i = cell i, nx <= MAXT,
// nx even) of plane k and marches up the rows j0 .. j0+seg-1.  Rows of the
// primitive state go through a ring of three LDS row buffers, filled by
// direct global-to-LDS copies (global_load_lds_dwordx4, no registers on
// the way) one row ahead.  For row j each thread
//   - reconstructs its own cell in x (PLM from the ring, periodic
//     neighbours), solves the x interface i-1/2 with its left neighbour's
//     right-face state from LDS, and takes the flux at i+1/2 from its right
//     neighbour through LDS;
//   - reconstructs cell (i, j+1) in y from rows j, j+1, j+2 of the ring and
//     solves the y interface j+1/2; the flux at j-1/2 and the right-face
//     state of cell (i, j) are carried from the previous row;
//   - adds the x then the y flux difference to U and writes the face data
//     of its faces and C_dt's x+y part.
// Every cell is reconstructed once per direction and every interface solved
// once (plus, per segment, the y interface j0-1/2 and two y
// reconstructions), where the tile kernel reconstructs each cell twice per
// direction and solves its halo interfaces again.  Two workgroup barriers
// per row.  The arithmetic per interface and per cell is the tile kernel's,
// so the results are identical.
// LDS (dynamic): 3 * NV * blockDim.x ring values and (NV + 7) * blockDim.x
// values of R.
template <int STAGE, typename R, int S, int E, int V, int MAXT>
__global__ void __launch_bounds__(MAXT, 2)
k_flux_xy_rows(Geo g, StageIn in, Cons U, FaceData fd, double* cdt_xy, int seg, DevScalars* sc)
{
  // rows of the state as they are stored: FP64, or float for the FP32-flux
  // path's second stage; CPL cells per 16-byte lane copy
  using RT = std::conditional_t<(STAGE == 2 && sizeof(R) == 4), float, double>;
  constexpr int CPL = 16 / int(sizeof(RT)), CCH = 64 * CPL;
  extern __shared__ __align__(16) unsigned char xrow_lds[];
  __shared__ double red[MAXT / 64 + 1];
  const int nt = blockDim.x;
  RT* ring = reinterpret_cast<RT*>(xrow_lds);          // [3][NV][nt] rows of the state
  R* xp = reinterpret_cast<R*>(ring + 3 * NV * nt);   // [NV][nt] right-face states, B_x: the face value
  R* xf = xp + NV * nt;                               // [7][nt] flux at i-1/2: rho mx my mz E p cmax

  const int i = threadIdx.x;
  const int lane = i & 63;
  const int wave = __builtin_amdgcn_readfirstlane(i >> 6);
  const int nwave = nt >> 6;
  const bool act = i < g.nx;
  const int ic = act ? i : g.nx - 1;
  const int il = ic == 0 ? g.nx - 1 : ic - 1;
  const int ir = ic == g.nx - 1 ? 0 : ic + 1;
  const dim3 tb = xcd_tile();
  const int j0 = tb.x * seg, k = tb.y;
  const int j1 = min(j0 + seg, g.ny);
  sc = scal(sc, g, k);
  // 32-bit byte offsets of FP64 fields (the launcher keeps this kernel to
  // boxes under 4 GiB per field): the loop's addresses stay one VGPR each
  const unsigned plane8 = (unsigned)k * (unsigned)g.ny * (unsigned)g.nx * 8u;
  const unsigned row8 = (unsigned)g.nx * 8u;
  const double dt = sc->dt;
  const double dtdx = dt / g.dx[0], dtdy = dt / g.dx[1];
  const R gam = R(g.gamma), gm1 = R(g.gmm1);
  double mach = 0.0;

  auto off8 = [&](int jj, int ii) { return plane8 + (unsigned)jj * row8 + (unsigned)ii * 8u; };
  auto load = [&](int jj, R* v) {
    const unsigned o = off8(wrap_any(jj, g.ny), ic);
#pragma unroll
    for (int q = 0; q < NV; q++)
      v[q] = sizeof(RT) == 4 ? R(at_byte(reinterpret_cast<const float*>(in.V[q]), o >> 1))
                             : R(at_byte(in.V[q], o));
  };
  // face data: FP64, or float in the first half of the buffer
  auto fstore = [&](double* a, unsigned o, R v) {
    if constexpr (sizeof(R) == 4) at_byte(reinterpret_cast<float*>(a), o >> 1) = v;
    else at_byte(a, o) = v;
  };
  // copy row jj of every variable into ring slot sl: 16 bytes (CPL cells)
  // per lane, chunks of CCH cells dealt to the waves in turn
  const int nchunk = (g.nx + CCH - 1) / CCH;
  auto stage_row = [&](int jj, int sl) {
    const unsigned o = off8(wrap_any(jj, g.ny), 0) / (8u / sizeof(RT));   // byte offset of the row
    for (int n = wave; n < NV * nchunk; n += nwave) {
      const int q = n / nchunk, c = n - q * nchunk;
      const int cell = c * CCH + CPL * lane;
      if (cell < g.nx)
        __builtin_amdgcn_global_load_lds(
            (const void*)(reinterpret_cast<const char*>(in.V[q]) + (o + (unsigned)cell * (unsigned)sizeof(RT))),
            (__attribute__((address_space(3))) void*)(ring + (sl * NV + q) * nt + c * CCH), 16, 0, 0);
    }
  };
  // wait for every vector memory access of this thread (the ring copies
  // among them) before a barrier that publishes the copies
  auto wait_vm = [] { __builtin_amdgcn_s_waitcnt(0x0070); };   // vmcnt(0) lgkmcnt(0)
  auto rv = [&](int sl, int q, int ii) -> R { return R(ring[(sl * NV + q) * nt + ii]); };
  // velocity slopes of the y reconstruction of cell (i, jj)
  auto store_dvy = [&](int jj, const R* vp, const R* vm) {
    if constexpr (E != FACE_BFLUX) {
      const unsigned o = off8(jj, ic);
      fstore(fd.dv[1][0], o, vp[VX1] - vm[VX1]);
      fstore(fd.dv[1][1], o, vp[VX2] - vm[VX2]);
      fstore(fd.dv[1][2], o, vp[VX3] - vm[VX3]);
    }
  };

  stage_row(j0, 0);
  stage_row(j0 + 1, 1);
  // the y right-face state of cell (i, j0) and the y flux at j0-1/2
  R vpy[NV], fyb[6], cmyb;
  {
    R a[NV], b[NV], w0[NV], w1[NV], vpb[NV], vmb[NV], vm0[NV];
    load(j0 - 2, a);
    load(j0 - 1, b);
    load(j0, w0);
    load(j0 + 1, w1);
    plm_cell(a, b, w0, vpb, vmb);      // cell (i, j0-1)
    plm_cell(b, w0, w1, vpy, vm0);     // cell (i, j0)
    if (act) store_dvy(j0, vpy, vm0);
    R vL[NV], vR[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
      vL[q] = vpb[q];
      vR[q] = vm0[q];
    }
    const R bface = R(at_byte(in.b[1], off8(wrap_any(j0 - 1, g.ny), ic)));
    vL[BX2] = bface;
    vR[BX2] = bface;
    RiemannOut<R> r;
    riemann<S, 1, E == FACE_BFLUX>(vL, vR, gam, gm1, r, g.dims == 2);
    fyb[0] = r.f[RHO]; fyb[1] = r.f[VX1]; fyb[2] = r.f[VX2];
    fyb[3] = r.f[VX3]; fyb[4] = r.f[ENG]; fyb[5] = r.press;
    cmyb = r.cmax;
    if (act && in_sweep(g, 1, ic, j0 - 1, k)) mach = double(r.mach);
  }
  wait_vm();
  __syncthreads();

  // FP64: the face values of a row are loaded during the row before, ahead
  // of its stores (vector memory completes in order: loaded after them, they
  // would wait for them), and are in place by that row's last barrier
  // (1.33 -> 1.30 ms at 256^3; the FP32 kernel measured slower so and loads
  // them at the top of the row)
  constexpr bool kPrefetch = sizeof(R) == sizeof(double);
  R bfx_n = R(0), bfy_n = R(0);
  if (kPrefetch) {
    bfx_n = R(at_byte(in.b[0], off8(j0, il)));
    bfy_n = R(at_byte(in.b[1], off8(j0, ic)));
  }
  int s0 = 0, s1 = 1, s2 = 2;   // ring slots of rows j, j+1, j+2
  for (int j = j0; j < j1; j++) {
    const unsigned c = off8(j, ic);
    const unsigned cf = off8(j, il);   // owner cell of the x face i-1/2
    const R bfx = kPrefetch ? bfx_n : R(at_byte(in.b[0], cf));
    const R bfy = kPrefetch ? bfy_n : R(at_byte(in.b[1], c));

    // x reconstruction of cell (i, j)
    double u[5];
    R vmx[NV];
    {
      R w0[NV], a[NV], b[NV], vpx[NV];
#pragma unroll
      for (int q = 0; q < NV; q++) w0[q] = rv(s0, q, ic);
      if (STAGE == 1) {   // U0 from the FP64 state
        double v[NV];
#pragma unroll
        for (int q = 0; q < NV; q++) v[q] = double(ring[(s0 * NV + q) * nt + ic]);
        u[0] = v[RHO];
        u[1] = v[RHO] * v[VX1];
        u[2] = v[RHO] * v[VX2];
        u[3] = v[RHO] * v[VX3];
        u[4] = total_energy(v, g.gmm1);
      }
#pragma unroll
      for (int q = 0; q < XROW_NV; q++) {
        a[xrow_var(q)] = rv(s0, xrow_var(q), il);
        b[xrow_var(q)] = rv(s0, xrow_var(q), ir);
      }
      a[BX1] = w0[BX1];   // not used: the x interfaces take the face value
      b[BX1] = w0[BX1];
      plm_cell(a, w0, b, vpx, vmx);
      // (every thread writes its slot: the face value's load then stays
      // at the top of the row instead of sinking into a branch)
#pragma unroll
      for (int q = 0; q < NV; q++) xp[q * nt + i] = q == BX1 ? bfx : vpx[q];
      if (act) {
        if constexpr (E != FACE_BFLUX) {
          fstore(fd.dv[0][0], c, vpx[VX1] - vmx[VX1]);
          fstore(fd.dv[0][1], c, vpx[VX2] - vmx[VX2]);
          fstore(fd.dv[0][2], c, vpx[VX3] - vmx[VX3]);
        }
      }
    }
    __syncthreads();
    // row j+2 into the slot of row j-1 (last read before the barrier)
    if (j + 1 < j1) {
      stage_row(j + 2, s2);
      if (kPrefetch) {
        bfx_n = R(at_byte(in.b[0], off8(j + 1, il)));
        bfy_n = R(at_byte(in.b[1], off8(j + 1, ic)));
      }
    }

    // x interface i-1/2
    R fxl[6], cmxl, xsl, xsr, xe[3];
    {
      R vL[NV], vR[NV];
#pragma unroll
      for (int q = 0; q < NV; q++) vL[q] = xp[q * nt + il];
#pragma unroll
      for (int q = 0; q < NV; q++) vR[q] = vmx[q];
      const R bface = xp[BX1 * nt + i];
      vL[BX1] = bface;
      vR[BX1] = bface;
      RiemannOut<R> r;
      riemann<S, 0, E == FACE_BFLUX>(vL, vR, gam, gm1, r, g.dims == 2);
      fxl[0] = r.f[RHO]; fxl[1] = r.f[VX1]; fxl[2] = r.f[VX2];
      fxl[3] = r.f[VX3]; fxl[4] = r.f[ENG]; fxl[5] = r.press;
      cmxl = r.cmax;
      if constexpr (E == FACE_BFLUX) {
        xe[0] = -r.f[BX2];   // ezi, eyi (ct_emf.c:131-133)
        xe[1] = r.f[BX3];
        xe[2] = r.f[RHO];
      } else {
        xsl = max_ge(R(0), -r.SL);
        xsr = max_ge(R(0), r.SR);
      }
      if (act) {
#pragma unroll
        for (int q = 0; q < 6; q++) xf[q * nt + i] = fxl[q];
        xf[6 * nt + i] = cmxl;
        const double rm = double(r.mach);
        if (in_sweep(g, 0, ic - 1, j, k)) mach = rm > mach ? rm : mach;
      }
    }
    wait_vm();
    __syncthreads();   // the copy of row j+2 is in place
    double U0[5];
    if (STAGE == 2) {
#pragma unroll
      for (int q = 0; q < 5; q++) U0[q] = at_byte(U.u[q], c);
    }
    if (act) {
      if constexpr (E == FACE_BFLUX) {
        at_byte(fd.SL[0], cf) = xe[0];
        at_byte(fd.SR[0], cf) = xe[1];
        at_byte(fd.dv[0][0], cf) = xe[2];
      } else {
        fstore(fd.SL[0], cf, xsl);
        fstore(fd.SR[0], cf, xsr);
      }
    }
    double rx[5], cdt = 0.0;
#pragma unroll
    for (int q = 0; q < 5; q++) rx[q] = -dtdx * (double(xf[q * nt + ir]) - double(fxl[q]));
    rx[1] -= dtdx * (double(xf[5 * nt + ir]) - double(fxl[5]));
    if (STAGE == 1) cdt += 0.5 * (double(cmxl) + double(xf[6 * nt + ir])) * g.inv_dx[0];

    // y reconstruction of cell (i, j+1) and the y interface j+1/2
    R vpn[NV], vmn[NV];
    {
      R a[NV], b[NV], w2[NV];
#pragma unroll
      for (int q = 0; q < NV; q++) {
        a[q] = rv(s0, q, ic);
        b[q] = rv(s1, q, ic);
      }
      if (j + 1 < j1) {
#pragma unroll
        for (int q = 0; q < NV; q++) w2[q] = rv(s2, q, ic);
      } else {
        load(j + 2, w2);   // the segment's last row: row j+2 is not staged
      }
      plm_cell(a, b, w2, vpn, vmn);
    }
    if (act && j + 1 < j1) store_dvy(j + 1, vpn, vmn);
    {
      R vL[NV], vR[NV];
#pragma unroll
      for (int q = 0; q < NV; q++) {
        vL[q] = vpy[q];
        vR[q] = vmn[q];
      }
      vL[BX2] = bfy;
      vR[BX2] = bfy;
      RiemannOut<R> r;
      riemann<S, 1, E == FACE_BFLUX>(vL, vR, gam, gm1, r, g.dims == 2);
      const R fcur[6] = {r.f[RHO], r.f[VX1], r.f[VX2], r.f[VX3], r.f[ENG], r.press};
      if (act) {
        if constexpr (E == FACE_BFLUX) {
          at_byte(fd.SL[1], c) = r.f[BX1];    // ezj, exj (ct_emf.c:153-155)
          at_byte(fd.SR[1], c) = -r.f[BX3];
          at_byte(fd.dv[1][0], c) = r.f[RHO];
        } else {
          fstore(fd.SL[1], c, max_ge(R(0), -r.SL));
          fstore(fd.SR[1], c, max_ge(R(0), r.SR));
        }
        const double rm = double(r.mach);
        if (in_sweep(g, 1, ic, j, k)) mach = rm > mach ? rm : mach;
        double ry[5];
#pragma unroll
        for (int q = 0; q < 5; q++) ry[q] = -dtdy * (double(fcur[q]) - double(fyb[q]));
        ry[2] -= dtdy * (double(fcur[5]) - double(fyb[5]));
        if (STAGE == 2) {
#pragma unroll
          for (int q = 0; q < 5; q++) u[q] = U0[q];
        }
#pragma unroll
        for (int q = 0; q < 5; q++) u[q] += rx[q];
#pragma unroll
        for (int q = 0; q < 5; q++) u[q] += ry[q];
#pragma unroll
        for (int q = 0; q < 5; q++) at_byte(U.u[q], c) = u[q];
        if (STAGE == 1) {
          cdt += 0.5 * (double(cmyb) + double(r.cmax)) * g.inv_dx[1];
          at_byte(cdt_xy, c) = cdt;
        }
      }
#pragma unroll
      for (int q = 0; q < 6; q++) fyb[q] = fcur[q];
      cmyb = r.cmax;
    }
#pragma unroll
    for (int q = 0; q < NV; q++) vpy[q] = vpn[q];
    const int st = s0;
    s0 = s1;
    s1 = s2;
    s2 = st;
  }
  block_max(mach, &sc->mach_max, red);
}

// ---------------------------------------------------------------------
// z sweep, one thread per column marching over a z segment
// ---------------------------------------------------------------------

template <int STAGE, typename R, int S, int E, int V = 0>
__global__ void __launch_bounds__(Z_TX * Z_TY)
k_flux_z(Geo g, StageIn in, Cons U, FaceData fd, const double* cdt_xy, int seg, DevScalars* sc)
{
  __shared__ double red[(Z_TX * Z_TY) / 64 + 1];
  const int i = blockIdx.x * Z_TX + threadIdx.x;
  const int j = blockIdx.y * Z_TY + threadIdx.y;
  const ZSeg zs = zseg(g, blockIdx.z, seg);   // (a patch of a stack: z periodic within it)
  const int k0 = zs.k0, k1 = zs.k1;
  sc = scal(sc, g, k0);
  double mach = 0.0, cdt_max = 0.0;
  if (i < g.nx && j < g.ny) {
    // 32-bit byte offsets (the launcher keeps the z march to fields under
    // 4 GiB): the column's addresses stay one VGPR each
    const unsigned col8 = ((unsigned)j * (unsigned)g.nx + (unsigned)i) * 8u;
    const unsigned pl8 = (unsigned)g.nx * (unsigned)g.ny * 8u;
    const double dt = sc->dt;
    const double dtdz = dt / g.dx[2];
    const R gam = R(g.gamma), gm1 = R(g.gmm1);
    auto load = [&](int kk, R* v) {
      const unsigned o = (unsigned)zwrap(g, zs.kb, kk) * pl8 + col8;
#pragma unroll
      for (int q = 0; q < NV; q++)   // the FP32 path's second stage reads its input as float
        v[q] = (STAGE == 2 && sizeof(R) == 4) ? R(at_byte(reinterpret_cast<const float*>(in.V[q]), o >> 1))
                                              : R(at_byte(in.V[q], o));
    };
    auto fstore = [&](double* a, unsigned o, R v) {
      if constexpr (sizeof(R) == 4) at_byte(reinterpret_cast<float*>(a), o >> 1) = v;
      else at_byte(a, o) = v;
    };
    // vb = v(k), vc = v(k+1); vp_k: right-face state of cell k, dv_k its
    // velocity slopes
    R vb[NV], vc[NV], vp_k[NV], dv_k[3];
    R fprev[6], cmprev;
    {
      R va[NV], vm_k[NV], vp_n[NV], vm_n[NV];
      load(k0 - 2, va);
      load(k0 - 1, vb);
      load(k0, vc);
      plm_cell(va, vb, vc, vp_k, vm_k);      // cell k0-1
#pragma unroll
      for (int q = 0; q < NV; q++) va[q] = vb[q];
#pragma unroll
      for (int q = 0; q < NV; q++) vb[q] = vc[q];
      load(k0 + 1, vc);
      plm_cell(va, vb, vc, vp_n, vm_n);      // cell k0
      R vL[NV], vR[NV];
#pragma unroll
      for (int q = 0; q < NV; q++) {
        vL[q] = vp_k[q];
        vR[q] = vm_n[q];
      }
      const R bface = R(at_byte(in.b[2], (unsigned)zwrap(g, zs.kb, k0 - 1) * pl8 + col8));
      vL[BX3] = bface;
      vR[BX3] = bface;
      RiemannOut<R> r;
      riemann<S, 2, false>(vL, vR, gam, gm1, r);
      if (in_sweep(g, 2, i, j, k0 - 1)) mach = double(r.mach);
      fprev[0] = r.f[RHO]; fprev[1] = r.f[VX1]; fprev[2] = r.f[VX2];
      fprev[3] = r.f[VX3]; fprev[4] = r.f[ENG]; fprev[5] = r.press;
      cmprev = r.cmax;
#pragma unroll
      for (int q = 0; q < NV; q++) vp_k[q] = vp_n[q];   // cell k0
#pragma unroll
      for (int q = 0; q < 3; q++) dv_k[q] = vp_n[VX1 + q] - vm_n[VX1 + q];
    }
    // march: vb = v(k), vc = v(k+1)
    for (int k = k0; k < k1; k++) {
      const unsigned c = (unsigned)k * pl8 + col8;
      R vn[NV], vp_n[NV], vm_n[NV];
      load(k + 2, vn);
      const R bface = R(at_byte(in.b[2], c));
      // U and C_dt(x+y) of cell k read here, ahead of this plane's stores
      // (which may alias them as far as the compiler knows): their latency
      // then passes under the solve
      double u[5], cdt = 0.0;
#pragma unroll
      for (int q = 0; q < 5; q++) u[q] = at_byte(U.u[q], c);
      if (STAGE == 1) cdt = at_byte(cdt_xy, c);
      plm_cell(vb, vc, vn, vp_n, vm_n);     // cell k+1
      if constexpr (E != FACE_BFLUX) {
        fstore(fd.dv[2][0], c, dv_k[0]);
        fstore(fd.dv[2][1], c, dv_k[1]);
        fstore(fd.dv[2][2], c, dv_k[2]);
      }
      R vL[NV], vR[NV];
#pragma unroll
      for (int q = 0; q < NV; q++) {
        vL[q] = vp_k[q];
        vR[q] = vm_n[q];
      }
      vL[BX3] = bface;
      vR[BX3] = bface;
      RiemannOut<R> r;
      riemann<S, 2, E == FACE_BFLUX>(vL, vR, gam, gm1, r);
      const double rm = double(r.mach);
      if (in_sweep(g, 2, i, j, k)) mach = rm > mach ? rm : mach;
      const R fcur[6] = {r.f[RHO], r.f[VX1], r.f[VX2], r.f[VX3], r.f[ENG], r.press};
      double rr[5];
#pragma unroll
      for (int q = 0; q < 5; q++) rr[q] = -dtdz * (double(fcur[q]) - double(fprev[q]));
      rr[3] -= dtdz * (double(fcur[5]) - double(fprev[5]));
#pragma unroll
      for (int q = 0; q < 5; q++) u[q] += rr[q];
#pragma unroll
      for (int q = 0; q < 5; q++) at_byte(U.u[q], c) = u[q];
      if constexpr (E == FACE_BFLUX) {
        at_byte(fd.SL[2], c) = -r.f[BX1];    // eyk, exk (ct_emf.c:175-176)
        at_byte(fd.SR[2], c) = r.f[BX2];
        at_byte(fd.dv[2][0], c) = r.f[RHO];
      } else {
        fstore(fd.SL[2], c, max_ge(R(0), -r.SL));
        fstore(fd.SR[2], c, max_ge(R(0), r.SR));
      }
      if (STAGE == 1) {
        cdt += 0.5 * (double(cmprev) + double(r.cmax)) * g.inv_dx[2];
        if (in_dom(g, i, j, k)) cdt_max = cdt > cdt_max ? cdt : cdt_max;
      }
#pragma unroll
      for (int q = 0; q < 6; q++) fprev[q] = fcur[q];
      cmprev = r.cmax;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        vp_k[q] = vp_n[q];
        vb[q] = vc[q];
        vc[q] = vn[q];
      }
#pragma unroll
      for (int q = 0; q < 3; q++) dv_k[q] = vp_n[VX1 + q] - vm_n[VX1 + q];
    }
  }
  block_max(mach, &sc->mach_max, red);
  if (STAGE == 1) {
    __syncthreads();
    block_max(cdt_max, &sc->invdt_max, red);
  }
}

// 2D problems: C_dt's maximum directly from the x+y part.
// (a template so that each translation unit that launches it has its own)
template <int = 0>
__global__ void k_cdt_max2d(Geo g, const double* cdt_xy, DevScalars* sc)
{
  __shared__ double red[4];
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = c < g.n ? cdt_xy[c] : 0.0;
  block_max(v, &sc->invdt_max, red);
}

template <typename R, int S, int E = FACE_UCT_HLL, int V = 0>
static void launch_flux_fused_t(const Geo& g, int stage, const StageIn& in, const Cons& U,
                                const FaceData& fd, double* cdt_xy, DevScalars* sc,
                                hipStream_t s, KernelTimer* kt)
{
  if (kt) kt->begin(KT_FLUX_XY, s);
  bool rows = false;
  {
    // the row kernel: rows of up to 256 cells, an even number (FP64) or a
    // multiple of 4 (FP32 flux: the second stage's float rows), fields under
    // 4 GiB (Roe's FP64 row kernel spills about 50 VGPRs and still beats the
    // tiles: 2.37 -> 2.07 ms at 256^3)
    const int nt = (g.nx + 63) / 64 * 64;
    rows = flux_rows_ok(g, (int)sizeof(R), S);
    if (rows) {
      const int seg = flux_xy_seg(g);
      const dim3 grid((g.ny + seg - 1) / seg, g.nz), block(nt);
      // ring rows: FP64, except the FP32-flux path's float second-stage input
      const size_t rt = (stage == 2 && sizeof(R) == 4) ? 4 : 8;
      const size_t lds = 3 * NV * nt * rt + (NV + 7) * nt * sizeof(R);
      if (stage == 1)
        hipLaunchKernelGGL((k_flux_xy_rows<1, R, S, E, V, 256>), grid, block, lds, s, g, in, U, fd, cdt_xy, seg, sc);
      else
        hipLaunchKernelGGL((k_flux_xy_rows<2, R, S, E, V, 256>), grid, block, lds, s, g, in, U, fd, cdt_xy, seg, sc);
    }
  }
  if (rows) {
  } else if constexpr (sizeof(R) == sizeof(double)) {
    constexpr int TX = 32;
    const dim3 bxy(TX, XY_TY), gxy((g.nx + TX - 2) / (TX - 1), (g.ny + XY_TY - 2) / (XY_TY - 1), g.nz);
    if (stage == 1) hipLaunchKernelGGL((k_flux_xy<1, R, TX, S, E, V>), gxy, bxy, 0, s, g, in, U, fd, cdt_xy, sc);
    else hipLaunchKernelGGL((k_flux_xy<2, R, TX, S, E, V>), gxy, bxy, 0, s, g, in, U, fd, cdt_xy, sc);
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
    const int pz = patch_depth(g);
    int seg = pz >= 64 ? 32 : (pz >= 16 ? 8 : pz);
    if (seg == 32 && xy * zblocks(g, seg) < 512) seg = 16;
    const dim3 bz(Z_TX, Z_TY), gz((g.nx + Z_TX - 1) / Z_TX, (g.ny + Z_TY - 1) / Z_TY,
                                  zblocks(g, seg));
    if (kt) kt->begin(KT_FLUX_Z, s);
    if (stage == 1)
      hipLaunchKernelGGL((k_flux_z<1, R, S, E, V>), gz, bz, 0, s, g, in, U, fd, cdt_xy, seg, sc);
    else
      hipLaunchKernelGGL((k_flux_z<2, R, S, E, V>), gz, bz, 0, s, g, in, U, fd, cdt_xy, seg, sc);
    if (kt) kt->end(KT_FLUX_Z, s);
  } else if (stage == 1) {
    hipLaunchKernelGGL(k_cdt_max2d<>, dim3((unsigned)((g.n + 255) / 256)), dim3(256), 0, s, g,
                       cdt_xy, sc);
  }
}

}  // namespace pgpu
