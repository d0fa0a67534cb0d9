// Cell-level bodies of the modular stage kernels (stage_kernels.hip), shared
// with the batched-patch kernels (batch_kernels.hip).  Each function does the
// work of one cell p of a periodic box described by Geo; see
// stage_kernels.hip for the stage structure and the reference lines each
// step follows.
#pragma once
#include <hip/hip_runtime.h>
#include "mhd_device.hpp"
#include "stage_kernels.hpp"

namespace pgpu {

// ---------------------------------------------------------------------
// indexing helpers
// ---------------------------------------------------------------------

struct Cell {
  int i, j, k;
  long c;
};

// cell of this thread in the (64 x 4 x nz) launch grid of a box
__device__ __forceinline__ bool cell_of(const Geo& g, Cell& p)
{
  p.i = blockIdx.x * blockDim.x + threadIdx.x;
  p.j = blockIdx.y * blockDim.y + threadIdx.y;
  p.k = blockIdx.z;
  if (p.i >= g.nx || p.j >= g.ny) return false;
  p.c = ((long)p.k * g.ny + p.j) * g.nx + p.i;
  return true;
}

// neighbour of cell c at offset o (|o| <= 2) along direction D, periodic
template <int D>
__device__ __forceinline__ long nb(const Geo& g, const Cell& p, int o)
{
  const int q = (D == 0 ? p.i : D == 1 ? p.j : p.k);
  const int nq = (D == 0 ? g.nx : D == 1 ? g.ny : g.nz);
  const long s = (D == 0 ? 1L : D == 1 ? (long)g.nx : (long)g.nx * g.ny);
  int qq = q + o;
  if (qq < 0) qq += nq;
  else if (qq >= nq) qq -= nq;
  return p.c + (long)(qq - q) * s;
}

// ---------------------------------------------------------------------
// sweep
// ---------------------------------------------------------------------

// PLM + HLLD at the interface q+1/2 of cell p along D; returns the Mach
// estimate of that interface.
template <int D>
__device__ __forceinline__ double sweep_cell(const Geo& g, const Cell& p, const StageIn& in,
                                             const SweepOut& out)
{
  constexpr int BN = BX1 + Dir<D>::n;
  const long cm1 = nb<D>(g, p, -1), cp1 = nb<D>(g, p, 1), cp2 = nb<D>(g, p, 2);
  double vL[NV], vR[NV], dvc[3];
#pragma unroll
  for (int nv = 0; nv < NV; nv++) {
    const double* a = in.V[nv];
    const double v_m1 = a[cm1], v_0 = a[p.c], v_p1 = a[cp1], v_p2 = a[cp2];
    const double s0 = mc_slope(v_m1, v_0, v_p1);
    const double s1 = mc_slope(v_0, v_p1, v_p2);
    const double vp0 = v_0 + s0 * 0.5;   // cell q, right face
    const double vm0 = v_0 - s0 * 0.5;   // cell q, left face
    const double vm1 = v_p1 - s1 * 0.5;  // cell q+1, left face
    vL[nv] = vp0;
    vR[nv] = vm1;
    if (nv >= VX1 && nv <= VX3) dvc[nv - VX1] = vp0 - vm0;
  }
  const double bface = in.b[D][p.c];
  vL[BN] = bface;
  vR[BN] = bface;

  RiemannOut<double> r;
  hlld<D, false>(vL, vR, g.gamma, g.gmm1, r);

  out.F[0][p.c] = r.f[RHO];
  out.F[1][p.c] = r.f[VX1];
  out.F[2][p.c] = r.f[VX2];
  out.F[3][p.c] = r.f[VX3];
  out.F[4][p.c] = r.f[ENG];
  out.P[p.c] = r.press;
  out.SL[p.c] = max_ge(0.0, -r.SL);
  out.SR[p.c] = max_ge(0.0, r.SR);
  out.cm[p.c] = r.cmax;
  out.dv[0][p.c] = dvc[0];
  out.dv[1][p.c] = dvc[1];
  out.dv[2][p.c] = dvc[2];
  return r.mach;
}

// ---------------------------------------------------------------------
// update
// ---------------------------------------------------------------------

template <int D>
__device__ __forceinline__ void add_dir(const Geo& g, const Cell& p, const SweepOut& s,
                                        double dt, double* u, double& cdt, bool stage1)
{
  const long cl = nb<D>(g, p, -1);
  const double dtdx = dt / g.dx[D];
  double r[5];
#pragma unroll
  for (int q = 0; q < 5; q++) r[q] = -dtdx * (s.F[q][p.c] - s.F[q][cl]);
  r[1 + D] -= dtdx * (s.P[p.c] - s.P[cl]);
#pragma unroll
  for (int q = 0; q < 5; q++) u[q] += r[q];
  if (stage1) cdt += 0.5 * (s.cm[cl] + s.cm[p.c]) * g.inv_dx[D];
}

// Conservative update of cell p in direction order; FROM_V: the base state
// is PrimToCons of the stage input (stage 1 of AdvanceStep, rk_step.c:93),
// otherwise it is read from U.  CDT: returns C_dt of the cell
// (g_intStage == 1), else 0.
template <bool FROM_V, bool CDT>
__device__ __forceinline__ double update_cell(const Geo& g, const Cell& p, const StageIn& in,
                                              const Sweeps& sw, const Cons& U, double dt)
{
  double cdt = 0.0;
  double u[5];
  if (FROM_V) {
    double v[NV];
#pragma unroll
    for (int nv = 0; nv < NV; nv++) v[nv] = in.V[nv][p.c];
    u[0] = v[RHO];
    u[1] = v[RHO] * v[VX1];
    u[2] = v[RHO] * v[VX2];
    u[3] = v[RHO] * v[VX3];
    u[4] = total_energy(v, g.gmm1);
  } else {
#pragma unroll
    for (int q = 0; q < 5; q++) u[q] = U.u[q][p.c];
  }
  add_dir<0>(g, p, sw.d[0], dt, u, cdt, CDT);
  add_dir<1>(g, p, sw.d[1], dt, u, cdt, CDT);
  if (g.dims == 3) add_dir<2>(g, p, sw.d[2], dt, u, cdt, CDT);
#pragma unroll
  for (int q = 0; q < 5; q++) U.u[q][p.c] = u[q];
  return cdt;
}

// ---------------------------------------------------------------------
// emf
// ---------------------------------------------------------------------

// slope along axis A of a face array at the cell at offsets (oi, oj, ok)
template <int A>
__device__ __forceinline__ double face_slope(const Geo& g, const double* f, const Cell& p,
                                             int oi, int oj, int ok)
{
  Cell q = p;
  q.i = p.i + oi; if (q.i >= g.nx) q.i -= g.nx;
  q.j = p.j + oj; if (q.j >= g.ny) q.j -= g.ny;
  q.k = p.k + ok; if (q.k >= g.nz) q.k -= g.nz;
  q.c = ((long)q.k * g.ny + q.j) * g.nx + q.i;
  const long cp = nb<A>(g, q, 1), cm = nb<A>(g, q, -1);
  const double f0 = f[q.c];
  return mc_lim2(f[cp] - f0, f0 - f[cm]);
}

// UCT_HLL EMFs of the three edges owned by cell p.
__device__ __forceinline__ void emf_cell(const Geo& g, const Cell& p, const StageIn& in,
                                         const Sweeps& sw, const EdgeOut& e)
{
  const long c = p.c;
  const long cip = nb<0>(g, p, 1), cjp = nb<1>(g, p, 1);
  const long cipjp = cip + (cjp - c);
  const double* vx = in.V[VX1];
  const double* vy = in.V[VX2];
  const double* vz = in.V[VX3];
  const double *bx = in.b[0], *by = in.b[1], *bz = in.b[2];
  const SweepOut &X = sw.d[0], &Y = sw.d[1], &Z = sw.d[2];

  // Ez at (i+1/2, j+1/2, k)
  {
    const double a_xp = max_ge(X.SR[c], X.SR[cjp]);
    const double a_xm = max_ge(X.SL[c], X.SL[cjp]);
    const double a_yp = max_ge(Y.SR[c], Y.SR[cip]);
    const double a_ym = max_ge(Y.SL[c], Y.SL[cip]);
    const double bS = bx[c] + 0.5 * face_slope<1>(g, bx, p, 0, 0, 0);
    const double bW = by[c] + 0.5 * face_slope<0>(g, by, p, 0, 0, 0);
    const double bN = bx[cjp] - 0.5 * face_slope<1>(g, bx, p, 0, 1, 0);
    const double bE = by[cip] - 0.5 * face_slope<0>(g, by, p, 1, 0, 0);
    const double eSW = (vy[c] + 0.5 * (X.dv[1][c] + Y.dv[1][c])) * bS
                     - (vx[c] + 0.5 * (X.dv[0][c] + Y.dv[0][c])) * bW;
    const double eSE = (vy[cip] - 0.5 * (X.dv[1][cip] - Y.dv[1][cip])) * bS
                     - (vx[cip] - 0.5 * (X.dv[0][cip] - Y.dv[0][cip])) * bE;
    const double eNE = (vy[cipjp] - 0.5 * (X.dv[1][cipjp] + Y.dv[1][cipjp])) * bN
                     - (vx[cipjp] - 0.5 * (X.dv[0][cipjp] + Y.dv[0][cipjp])) * bE;
    const double eNW = (vy[cjp] + 0.5 * (X.dv[1][cjp] - Y.dv[1][cjp])) * bN
                     - (vx[cjp] + 0.5 * (X.dv[0][cjp] - Y.dv[0][cjp])) * bW;
    e.ez[c] = uct_hll(a_xp, a_xm, a_yp, a_ym, eSW, eSE, eNE, eNW, bS, bN, bW, bE);
  }
  if (g.dims < 3) return;
  const long ckp = nb<2>(g, p, 1);
  const long cjpkp = cjp + (ckp - c);
  const long cipkp = cip + (ckp - c);
  // Ex at (i, j+1/2, k+1/2)
  {
    const double a_xp = max_ge(Y.SR[c], Y.SR[ckp]);
    const double a_xm = max_ge(Y.SL[c], Y.SL[ckp]);
    const double a_yp = max_ge(Z.SR[c], Z.SR[cjp]);
    const double a_ym = max_ge(Z.SL[c], Z.SL[cjp]);
    const double bS = by[c] + 0.5 * face_slope<2>(g, by, p, 0, 0, 0);
    const double bW = bz[c] + 0.5 * face_slope<1>(g, bz, p, 0, 0, 0);
    const double bN = by[ckp] - 0.5 * face_slope<2>(g, by, p, 0, 0, 1);
    const double bE = bz[cjp] - 0.5 * face_slope<1>(g, bz, p, 0, 1, 0);
    const double eSW = (vz[c] + 0.5 * (Y.dv[2][c] + Z.dv[2][c])) * bS
                     - (vy[c] + 0.5 * (Y.dv[1][c] + Z.dv[1][c])) * bW;
    const double eSE = (vz[cjp] - 0.5 * (Y.dv[2][cjp] - Z.dv[2][cjp])) * bS
                     - (vy[cjp] - 0.5 * (Y.dv[1][cjp] - Z.dv[1][cjp])) * bE;
    const double eNE = (vz[cjpkp] - 0.5 * (Y.dv[2][cjpkp] + Z.dv[2][cjpkp])) * bN
                     - (vy[cjpkp] - 0.5 * (Y.dv[1][cjpkp] + Z.dv[1][cjpkp])) * bE;
    const double eNW = (vz[ckp] + 0.5 * (Y.dv[2][ckp] - Z.dv[2][ckp])) * bN
                     - (vy[ckp] + 0.5 * (Y.dv[1][ckp] - Z.dv[1][ckp])) * bW;
    e.ex[c] = uct_hll(a_xp, a_xm, a_yp, a_ym, eSW, eSE, eNE, eNW, bS, bN, bW, bE);
  }
  // Ey at (i+1/2, j, k+1/2); stencil offsets as in ct_emf_average.c:339-342
  {
    const double a_xp = max_ge(Z.SR[c], Z.SR[cip]);
    const double a_xm = max_ge(Z.SL[c], Z.SL[cip]);
    const double a_yp = max_ge(X.SR[c], X.SR[ckp]);
    const double a_ym = max_ge(X.SL[c], X.SL[ckp]);
    const double bS = bz[c] + 0.5 * face_slope<0>(g, bz, p, 0, 0, 0);
    const double bW = bx[c] + 0.5 * face_slope<2>(g, bx, p, 0, 0, 0);
    const double bN = bz[cip] - 0.5 * face_slope<0>(g, bz, p, 1, 0, 0);
    const double bE = bx[ckp] - 0.5 * face_slope<2>(g, bx, p, 0, 0, 1);
    const double eSW = (vx[c] + 0.5 * (Z.dv[0][c] + X.dv[0][c])) * bS
                     - (vz[c] + 0.5 * (Z.dv[2][c] + X.dv[2][c])) * bW;
    const double eSE = (vx[ckp] - 0.5 * (Z.dv[0][ckp] - X.dv[0][ckp])) * bS
                     - (vz[ckp] - 0.5 * (Z.dv[2][ckp] - X.dv[2][ckp])) * bE;
    const double eNE = (vx[cipkp] - 0.5 * (Z.dv[0][cipkp] + X.dv[0][cipkp])) * bN
                     - (vz[cipkp] - 0.5 * (Z.dv[2][cipkp] + X.dv[2][cipkp])) * bE;
    const double eNW = (vx[cip] + 0.5 * (Z.dv[0][cip] - X.dv[0][cip])) * bN
                     - (vz[cip] + 0.5 * (Z.dv[2][cip] - X.dv[2][cip])) * bW;
    e.ey[c] = uct_hll(a_xp, a_xm, a_yp, a_ym, eSW, eSE, eNE, eNW, bS, bN, bW, bE);
  }
}

// ---------------------------------------------------------------------
// ct, c2p
// ---------------------------------------------------------------------

// CT update of the three faces owned by cell p; stage 2 blends with b0.
template <int STAGE>
__device__ __forceinline__ void ct_cell(const Geo& g, const Cell& p, const EdgeOut& e,
                                        const FaceIn& bin, const FaceIn& b0, const FaceOut& bout,
                                        double dt)
{
  const double dtx = dt / g.dx[0], dty = dt / g.dx[1], dtz = dt / g.dx[2];
  const long c = p.c;
  const long cim = nb<0>(g, p, -1), cjm = nb<1>(g, p, -1);
  double nbx, nby, nbz;
  if (g.dims == 3) {
    const long ckm = nb<2>(g, p, -1);
    const double rx = 0.0 - dty * (e.ez[c] - e.ez[cjm]) + dtz * (e.ey[c] - e.ey[ckm]);
    const double ry = dtx * (e.ez[c] - e.ez[cim]) - dtz * (e.ex[c] - e.ex[ckm]);
    const double rz = -dtx * (e.ey[c] - e.ey[cim]) + dty * (e.ex[c] - e.ex[cjm]);
    nbx = bin.b[0][c] + rx;
    nby = bin.b[1][c] + ry;
    nbz = bin.b[2][c] + rz;
  } else {
    const double rx = 0.0 - dty * (e.ez[c] - e.ez[cjm]);
    const double ry = dtx * (e.ez[c] - e.ez[cim]);
    nbx = bin.b[0][c] + rx;
    nby = bin.b[1][c] + ry;
    nbz = bin.b[2][c];
  }
  if (STAGE == 2) {
    nbx = 0.5 * b0.b[0][c] + 0.5 * nbx;
    nby = 0.5 * b0.b[1][c] + 0.5 * nby;
    nbz = 0.5 * b0.b[2][c] + 0.5 * nbz;
  }
  bout.b[0][c] = nbx;
  bout.b[1][c] = nby;
  bout.b[2][c] = nbz;
}

// Face average, stage-2 blend and ConsToPrim of cell p; returns the
// ConsToPrim flag (nonzero when a floor was applied).
template <int STAGE>
__device__ __forceinline__ unsigned char c2p_cell(const Geo& g, const Cell& p, const Cons& U,
                                                  const FaceIn& b, const StageIn& vn,
                                                  const Prims& Vout)
{
  const long c = p.c;
  double u[NV], v[NV];
  u[RHO] = U.u[0][c];
  u[VX1] = U.u[1][c];
  u[VX2] = U.u[2][c];
  u[VX3] = U.u[3][c];
  u[ENG] = U.u[4][c];
  if (STAGE == 2) {
    double v0[NV];
#pragma unroll
    for (int nv = 0; nv < NV; nv++) v0[nv] = vn.V[nv][c];
    const double e0 = total_energy(v0, g.gmm1);
    u[RHO] = 0.5 * v0[RHO] + 0.5 * u[RHO];
    u[VX1] = 0.5 * (v0[RHO] * v0[VX1]) + 0.5 * u[VX1];
    u[VX2] = 0.5 * (v0[RHO] * v0[VX2]) + 0.5 * u[VX2];
    u[VX3] = 0.5 * (v0[RHO] * v0[VX3]) + 0.5 * u[VX3];
    u[ENG] = 0.5 * e0 + 0.5 * u[ENG];
  }
  u[BX1] = 0.5 * (b.b[0][c] + b.b[0][nb<0>(g, p, -1)]);
  u[BX2] = 0.5 * (b.b[1][c] + b.b[1][nb<1>(g, p, -1)]);
  u[BX3] = (g.dims == 3) ? 0.5 * (b.b[2][c] + b.b[2][nb<2>(g, p, -1)]) : 0.0;
  const unsigned char f = cons2prim(u, v, g.gmm1);
#pragma unroll
  for (int nv = 0; nv < NV; nv++) Vout.v[nv][c] = v[nv];
  if (STAGE == 1) {
    U.u[0][c] = u[RHO];
    U.u[1][c] = u[VX1];
    U.u[2][c] = u[VX2];
    U.u[3][c] = u[VX3];
    U.u[4][c] = u[ENG];
  }
  return f;
}

}  // namespace pgpu
