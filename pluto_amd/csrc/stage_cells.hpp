// Cell-level bodies of the modular stage kernels (stage_kernels.hip), shared
// with the batched-patch kernels (batch_kernels.hip).  Each function does the
// work of one cell p of a periodic box described by Geo; see
// stage_kernels.hip for the stage structure and the reference lines each
// step follows.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

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

// PLM + the Riemann solver S at the interface q+1/2 of cell p along D;
// returns the Mach estimate of that interface.  Branch recorders (null in
// the production kernels): own takes this cell's limiter cases, the HLLD
// regions and the face-speed maxima, next the limiter cases of cell q+1.
template <int D, int S = SOLVER_HLLD>
__device__ __forceinline__ double sweep_cell(const Geo& g, const Cell& p, const StageIn& in,
                                             const SweepOut& out, BranchRec* own = nullptr,
                                             BranchRec* next = nullptr)
{
  constexpr int BN = BX1 + Dir<D>::n;
  const long cm1 = nb<D>(g, p, -1), cp1 = nb<D>(g, p, 1), cp2 = nb<D>(g, p, 2);
  double vL[NV], vR[NV], dvc[3];
#pragma unroll
  for (int nv = 0; nv < NV; nv++) {
    const double* a = in.V[nv];
    const double v_m1 = a[cm1], v_0 = a[p.c], v_p1 = a[cp1], v_p2 = a[cp2];
    const double s0 = mc_slope(v_m1, v_0, v_p1, own);
    const double s1 = mc_slope(v_0, v_p1, v_p2, next);
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
  riemann<S, D, false>(vL, vR, g.gamma, g.gmm1, r, g.dims == 2, own);

  out.F[0][p.c] = r.f[RHO];
  out.F[1][p.c] = r.f[VX1];
  out.F[2][p.c] = r.f[VX2];
  out.F[3][p.c] = r.f[VX3];
  out.F[4][p.c] = r.f[ENG];
  out.P[p.c] = r.press;
  out.SL[p.c] = max_ge(0.0, -r.SL, own);
  out.SR[p.c] = max_ge(0.0, r.SR, own);
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

// Fields read by the UCT_HLL EMF stencil, by number: the face speeds
// max(0, SR), max(0, -SL) of the x, y, z sweeps, the staggered field, the
// velocity and the PLM velocity slopes (sweep direction, component).
enum EmfField {
  F_XSR, F_XSL, F_YSR, F_YSL, F_ZSR, F_ZSL,
  F_BX, F_BY, F_BZ, F_VX, F_VY, F_VZ,
  F_XDV0, F_XDV1, F_XDV2, F_YDV0, F_YDV1, F_YDV2, F_ZDV0, F_ZDV1, F_ZDV2,
  F_EMF_COUNT
};

struct EmfFields {
  const double* a[F_EMF_COUNT];
};

__device__ __forceinline__ EmfFields emf_fields(const StageIn& in, const Sweeps& sw)
{
  EmfFields f;
  for (int d = 0; d < 3; d++) {
    f.a[F_XSR + 2 * d] = sw.d[d].SR;
    f.a[F_XSL + 2 * d] = sw.d[d].SL;
    f.a[F_BX + d] = in.b[d];
    f.a[F_VX + d] = in.V[VX1 + d];
    for (int c = 0; c < 3; c++) f.a[F_XDV0 + 3 * d + c] = sw.d[d].dv[c];
  }
  return f;
}

// MC_LIM2 slope along axis AX of face field A at the offset (I, J, K)
template <int A, int AX, int I, int J, int K, class L>
__device__ __forceinline__ double emf_slope(L& ld, BranchRec* rec = nullptr)
{
  const double f0 = ld.template get<A, I, J, K>();
  const double fp = ld.template get<A, I + (AX == 0), J + (AX == 1), K + (AX == 2)>();
  const double fm = ld.template get<A, I - (AX == 0), J - (AX == 1), K - (AX == 2)>();
  return mc_lim2(fp - f0, f0 - fm, rec);
}

// UCT_HLL EMFs of the three edges owned by a cell (ct_emf_average.c:179-352
// with the MC_LIM2 face slopes of ct_stag_slopes.c:51-106): Ez at
// (i+1/2, j+1/2, k), Ex at (i, j+1/2, k+1/2), Ey at (i+1/2, j, k+1/2).  The
// loader's get<A, I, J, K>() returns field A at the periodic offset (I, J, K)
// of the cell (offsets in [-1, 2]), so the per-cell and the z-marching
// kernels share this body.  SPLIT: scheduling barriers between the three
// edges, so the compiler does not issue the loads of all three at once (the
// z-marching kernel: 202 VGPRs and no spills at 2 waves per SIMD, 0.96 ms per
// 256^3 launch against 1.18 ms unsplit at 3 waves per SIMD).
template <bool SPLIT = false, class L>
__device__ __forceinline__ void emf_edges(L& ld, bool three_d, double& ez, double& ex, double& ey,
                                          BranchRec* rec = nullptr)
{
#define G(A, I, J, K) ld.template get<A, I, J, K>()
  // Ez at (i+1/2, j+1/2, k)
  {
    const double a_xp = max_ge(G(F_XSR, 0, 0, 0), G(F_XSR, 0, 1, 0), rec);
    const double a_xm = max_ge(G(F_XSL, 0, 0, 0), G(F_XSL, 0, 1, 0), rec);
    const double a_yp = max_ge(G(F_YSR, 0, 0, 0), G(F_YSR, 1, 0, 0), rec);
    const double a_ym = max_ge(G(F_YSL, 0, 0, 0), G(F_YSL, 1, 0, 0), rec);
    const double bS = G(F_BX, 0, 0, 0) + 0.5 * emf_slope<F_BX, 1, 0, 0, 0>(ld, rec);
    const double bW = G(F_BY, 0, 0, 0) + 0.5 * emf_slope<F_BY, 0, 0, 0, 0>(ld, rec);
    const double bN = G(F_BX, 0, 1, 0) - 0.5 * emf_slope<F_BX, 1, 0, 1, 0>(ld, rec);
    const double bE = G(F_BY, 1, 0, 0) - 0.5 * emf_slope<F_BY, 0, 1, 0, 0>(ld, rec);
    const double eSW = (G(F_VY, 0, 0, 0) + 0.5 * (G(F_XDV1, 0, 0, 0) + G(F_YDV1, 0, 0, 0))) * bS
                     - (G(F_VX, 0, 0, 0) + 0.5 * (G(F_XDV0, 0, 0, 0) + G(F_YDV0, 0, 0, 0))) * bW;
    const double eSE = (G(F_VY, 1, 0, 0) - 0.5 * (G(F_XDV1, 1, 0, 0) - G(F_YDV1, 1, 0, 0))) * bS
                     - (G(F_VX, 1, 0, 0) - 0.5 * (G(F_XDV0, 1, 0, 0) - G(F_YDV0, 1, 0, 0))) * bE;
    const double eNE = (G(F_VY, 1, 1, 0) - 0.5 * (G(F_XDV1, 1, 1, 0) + G(F_YDV1, 1, 1, 0))) * bN
                     - (G(F_VX, 1, 1, 0) - 0.5 * (G(F_XDV0, 1, 1, 0) + G(F_YDV0, 1, 1, 0))) * bE;
    const double eNW = (G(F_VY, 0, 1, 0) + 0.5 * (G(F_XDV1, 0, 1, 0) - G(F_YDV1, 0, 1, 0))) * bN
                     - (G(F_VX, 0, 1, 0) + 0.5 * (G(F_XDV0, 0, 1, 0) - G(F_YDV0, 0, 1, 0))) * bW;
    ez = uct_hll(a_xp, a_xm, a_yp, a_ym, eSW, eSE, eNE, eNW, bS, bN, bW, bE);
  }
  if (!three_d) return;
  if (SPLIT) __builtin_amdgcn_sched_barrier(0);
  // Ex at (i, j+1/2, k+1/2)
  {
    const double a_xp = max_ge(G(F_YSR, 0, 0, 0), G(F_YSR, 0, 0, 1), rec);
    const double a_xm = max_ge(G(F_YSL, 0, 0, 0), G(F_YSL, 0, 0, 1), rec);
    const double a_yp = max_ge(G(F_ZSR, 0, 0, 0), G(F_ZSR, 0, 1, 0), rec);
    const double a_ym = max_ge(G(F_ZSL, 0, 0, 0), G(F_ZSL, 0, 1, 0), rec);
    const double bS = G(F_BY, 0, 0, 0) + 0.5 * emf_slope<F_BY, 2, 0, 0, 0>(ld, rec);
    const double bW = G(F_BZ, 0, 0, 0) + 0.5 * emf_slope<F_BZ, 1, 0, 0, 0>(ld, rec);
    const double bN = G(F_BY, 0, 0, 1) - 0.5 * emf_slope<F_BY, 2, 0, 0, 1>(ld, rec);
    const double bE = G(F_BZ, 0, 1, 0) - 0.5 * emf_slope<F_BZ, 1, 0, 1, 0>(ld, rec);
    const double eSW = (G(F_VZ, 0, 0, 0) + 0.5 * (G(F_YDV2, 0, 0, 0) + G(F_ZDV2, 0, 0, 0))) * bS
                     - (G(F_VY, 0, 0, 0) + 0.5 * (G(F_YDV1, 0, 0, 0) + G(F_ZDV1, 0, 0, 0))) * bW;
    const double eSE = (G(F_VZ, 0, 1, 0) - 0.5 * (G(F_YDV2, 0, 1, 0) - G(F_ZDV2, 0, 1, 0))) * bS
                     - (G(F_VY, 0, 1, 0) - 0.5 * (G(F_YDV1, 0, 1, 0) - G(F_ZDV1, 0, 1, 0))) * bE;
    const double eNE = (G(F_VZ, 0, 1, 1) - 0.5 * (G(F_YDV2, 0, 1, 1) + G(F_ZDV2, 0, 1, 1))) * bN
                     - (G(F_VY, 0, 1, 1) - 0.5 * (G(F_YDV1, 0, 1, 1) + G(F_ZDV1, 0, 1, 1))) * bE;
    const double eNW = (G(F_VZ, 0, 0, 1) + 0.5 * (G(F_YDV2, 0, 0, 1) - G(F_ZDV2, 0, 0, 1))) * bN
                     - (G(F_VY, 0, 0, 1) + 0.5 * (G(F_YDV1, 0, 0, 1) - G(F_ZDV1, 0, 0, 1))) * bW;
    ex = uct_hll(a_xp, a_xm, a_yp, a_ym, eSW, eSE, eNE, eNW, bS, bN, bW, bE);
  }
  if (SPLIT) __builtin_amdgcn_sched_barrier(0);
  // Ey at (i+1/2, j, k+1/2); stencil offsets as in ct_emf_average.c:339-342
  {
    const double a_xp = max_ge(G(F_ZSR, 0, 0, 0), G(F_ZSR, 1, 0, 0), rec);
    const double a_xm = max_ge(G(F_ZSL, 0, 0, 0), G(F_ZSL, 1, 0, 0), rec);
    const double a_yp = max_ge(G(F_XSR, 0, 0, 0), G(F_XSR, 0, 0, 1), rec);
    const double a_ym = max_ge(G(F_XSL, 0, 0, 0), G(F_XSL, 0, 0, 1), rec);
    const double bS = G(F_BZ, 0, 0, 0) + 0.5 * emf_slope<F_BZ, 0, 0, 0, 0>(ld, rec);
    const double bW = G(F_BX, 0, 0, 0) + 0.5 * emf_slope<F_BX, 2, 0, 0, 0>(ld, rec);
    const double bN = G(F_BZ, 1, 0, 0) - 0.5 * emf_slope<F_BZ, 0, 1, 0, 0>(ld, rec);
    const double bE = G(F_BX, 0, 0, 1) - 0.5 * emf_slope<F_BX, 2, 0, 0, 1>(ld, rec);
    const double eSW = (G(F_VX, 0, 0, 0) + 0.5 * (G(F_ZDV0, 0, 0, 0) + G(F_XDV0, 0, 0, 0))) * bS
                     - (G(F_VZ, 0, 0, 0) + 0.5 * (G(F_ZDV2, 0, 0, 0) + G(F_XDV2, 0, 0, 0))) * bW;
    const double eSE = (G(F_VX, 0, 0, 1) - 0.5 * (G(F_ZDV0, 0, 0, 1) - G(F_XDV0, 0, 0, 1))) * bS
                     - (G(F_VZ, 0, 0, 1) - 0.5 * (G(F_ZDV2, 0, 0, 1) - G(F_XDV2, 0, 0, 1))) * bE;
    const double eNE = (G(F_VX, 1, 0, 1) - 0.5 * (G(F_ZDV0, 1, 0, 1) + G(F_XDV0, 1, 0, 1))) * bN
                     - (G(F_VZ, 1, 0, 1) - 0.5 * (G(F_ZDV2, 1, 0, 1) + G(F_XDV2, 1, 0, 1))) * bE;
    const double eNW = (G(F_VX, 1, 0, 0) + 0.5 * (G(F_ZDV0, 1, 0, 0) - G(F_XDV0, 1, 0, 0))) * bN
                     - (G(F_VZ, 1, 0, 0) + 0.5 * (G(F_ZDV2, 1, 0, 0) - G(F_XDV2, 1, 0, 0))) * bW;
    ey = uct_hll(a_xp, a_xm, a_yp, a_ym, eSW, eSE, eNE, eNW, bS, bN, bW, bE);
  }
#undef G
}

// periodic index q + o, |o| <= 2, of an axis of length n >= 2
__device__ __forceinline__ int wrap2(int q, int o, int n)
{
  int r = q + o;
  if (r < 0) r += n;
  else if (r >= n) r -= n;
  return r;
}

// Tile coordinates of this workgroup.  The dispatcher deals workgroups to
// the 8 XCDs round-robin by linear id, and each XCD has its own L2; the
// linear id is renumbered so that XCD x runs one contiguous range of tiles
// (x fastest, then y, then z).  Tiles that share a halo then run on the same
// XCD at about the same time, and the halo comes from its L2 instead of HBM.
// A bijection for any grid size, so it changes which workgroup computes a
// tile, never what is computed.  It pays in the overlapping FP64 x/y flux
// tile (about 2.5% at 256^3); in the z-marching EMF and CT kernels and the
// FP32 tile it measured slightly slower, so they keep the plain numbering.
__device__ __forceinline__ dim3 xcd_tile()
{
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned n = gx * gy * gridDim.z;
  const unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned x = lin & 7u, m = lin >> 3, q = n >> 3, r = n & 7u;
  const unsigned t = x * q + (x < r ? x : r) + m;
  const unsigned txy = t % (gx * gy);
  return dim3(txy % gx, txy / gx, t / (gx * gy));
}

// element at byte offset o of a field (32-bit offsets: fields under 4 GiB)
__device__ __forceinline__ const double& at_byte(const double* a, unsigned o)
{
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(a) + o);
}

__device__ __forceinline__ double& at_byte(double* a, unsigned o)
{
  return *reinterpret_cast<double*>(reinterpret_cast<char*>(a) + o);
}

__device__ __forceinline__ const float& at_byte(const float* a, unsigned o)
{
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(a) + o);
}

__device__ __forceinline__ float& at_byte(float* a, unsigned o)
{
  return *reinterpret_cast<float*>(reinterpret_cast<char*>(a) + o);
}

// Face speeds and velocity slopes of the EMF stencil are stored as FT: double,
// or float where the flux kernels computed them in FP32 (the first half of the
// same buffers).  Widening a float is exact, so the EMF sees the same values
// from half the bytes.
constexpr bool emf_face_field(int A) { return A <= F_ZSL || (A >= F_XDV0 && A <= F_ZDV2); }

// FT of the FP32-flux path's second stage: its input state (the stage-1
// output) is stored in FP32 too, so the velocities are float as well.
struct F32State {
  float v;
};

// whether field A is stored as float for face-data type FT
template <class FT>
constexpr bool emf_f32(int A)
{
  return (sizeof(FT) == 4 && emf_face_field(A)) ||
         (std::is_same<FT, F32State>::value && A >= F_VX && A <= F_VZ);
}

template <class FT, int A>
__device__ __forceinline__ double emf_field_at(const EmfFields& f, long c)
{
  if constexpr (emf_f32<FT>(A))
    return double(reinterpret_cast<const float*>(f.a[A])[c]);
  else
    return f.a[A][c];
}

// loader of emf_edges for one cell of a box (64-bit indices)
template <class FT = double>
struct EmfCellLoader {
  const EmfFields& f;
  const Geo& g;
  const Cell& p;
  template <int A, int I, int J, int K>
  __device__ __forceinline__ double get() const
  {
    const long o = (long)(wrap2(p.i, I, g.nx) - p.i) + (long)(wrap2(p.j, J, g.ny) - p.j) * g.nx +
                   (long)(wrap2(p.k, K, g.nz) - p.k) * ((long)g.nx * g.ny);
    return emf_field_at<FT, A>(f, p.c + o);
  }
};

// UCT_HLL EMFs of the three edges owned by cell p (face data stored as FT;
// rec: branch recorder, null in the production kernels).
template <class FT = double>
__device__ __forceinline__ void emf_cell(const Geo& g, const Cell& p, const StageIn& in,
                                         const Sweeps& sw, const EdgeOut& e,
                                         BranchRec* rec = nullptr)
{
  const EmfFields f = emf_fields(in, sw);
  EmfCellLoader<FT> ld{f, g, p};
  double ez, ex, ey;
  emf_edges(ld, g.dims == 3, ez, ex, ey, rec);
  e.ez[p.c] = ez;
  if (g.dims < 3) return;
  e.ex[p.c] = ex;
  e.ey[p.c] = ey;
}

// Edge EMFs of the ARITHMETIC, UCT0 and UCT_CONTACT averages at the three
// edges owned by cell p (CT_ComputeEMF, ct_emf.c:216-285;
// CT_EMF_ArithmeticAverage and CT_EMF_IntegrateToCorner,
// ct_emf_average.c:13-177).  The flux kernels store the face EMFs of
// CT_StoreUpwindEMF (ct_emf.c:131-187) in sw: x faces ezi = -F_x(By) in SL
// and eyi = F_x(Bz) in SR, y faces ezj = F_y(Bx) and exj = -F_y(Bz), z faces
// eyk = -F_z(Bx) and exk = F_z(By), and the face's mass flux in dv[0].  The
// cell EMF -v x B (CT_ComputeCenterEMF, ct_emf.c:348-388) comes from the
// stage input.  UCT_CONTACT: the reference adds its corner terms from a
// loop over cells in (k, j, i) order; an edge gets the two terms of its own
// cell, then the one of the next cell along the faster index, then that of
// the next cell along the slower index, and the sums below keep that order.
template <int MODE>
__device__ __forceinline__ void emf_faces_cell(const Geo& g, const Cell& p, const StageIn& in,
                                               const Sweeps& sw, const EdgeOut& e)
{
  // periodic cell at offset (di, dj, dk) in {0, 1}^3
  auto at = [&](int di, int dj, int dk) -> long {
    return p.c + (long)(wrap2(p.i, di, g.nx) - p.i) + (long)(wrap2(p.j, dj, g.ny) - p.j) * g.nx +
           (long)(wrap2(p.k, dk, g.nz) - p.k) * ((long)g.nx * g.ny);
  };
  // face EMF a (0: SL slot, 1: SR slot) of direction d
  auto F = [&](int d, int a, int di, int dj, int dk) -> double {
    return (a == 0 ? sw.d[d].SL : sw.d[d].SR)[at(di, dj, dk)];
  };
  // component a of the cell EMF
  auto E = [&](int a, int di, int dj, int dk) -> double {
    const long c = at(di, dj, dk);
    const double vx = in.V[VX1][c], vy = in.V[VX2][c], vz = in.V[VX3][c];
    const double bx = in.V[BX1][c], by = in.V[BX2][c], bz = in.V[BX3][c];
    return a == 0 ? vz * by - vy * bz : (a == 1 ? vx * bz - vz * bx : vy * bx - vx * by);
  };
  const bool three_d = g.dims == 3;
  double ez, ex = 0.0, ey = 0.0;

  if constexpr (MODE == EMF_UCT0) {
    // ct_emf.c:261-283: each face EMF doubled minus the mean of its two cells'
    auto U0 = [&](int d, int a, int c, int i, int j, int k, int di, int dj, int dk) -> double {
      return F(d, a, i, j, k) * 2.0 - 0.5 * (E(c, i, j, k) + E(c, i + di, j + dj, k + dk));
    };
    ez = 0.25 * (U0(0, 0, 2, 0, 0, 0, 1, 0, 0) + U0(0, 0, 2, 0, 1, 0, 1, 0, 0) +
                 U0(1, 0, 2, 0, 0, 0, 0, 1, 0) + U0(1, 0, 2, 1, 0, 0, 0, 1, 0));
    if (three_d) {
      ex = 0.25 * (U0(2, 1, 0, 0, 0, 0, 0, 0, 1) + U0(2, 1, 0, 0, 1, 0, 0, 0, 1) +
                   U0(1, 1, 0, 0, 0, 0, 0, 1, 0) + U0(1, 1, 0, 0, 0, 1, 0, 1, 0));
      ey = 0.25 * (U0(0, 1, 1, 0, 0, 0, 1, 0, 0) + U0(0, 1, 1, 0, 0, 1, 1, 0, 0) +
                   U0(2, 0, 1, 0, 0, 0, 0, 0, 1) + U0(2, 0, 1, 1, 0, 0, 0, 0, 1));
    }
  } else {
    const double w = MODE == EMF_ARITHMETIC ? 0.25 : 1.0;
    ez = w * (F(0, 0, 0, 0, 0) + F(0, 0, 0, 1, 0) + F(1, 0, 0, 0, 0) + F(1, 0, 1, 0, 0));
    if (three_d) {
      ex = w * (F(2, 1, 0, 0, 0) + F(2, 1, 0, 1, 0) + F(1, 1, 0, 0, 0) + F(1, 1, 0, 0, 1));
      ey = w * (F(0, 1, 0, 0, 0) + F(0, 1, 0, 0, 1) + F(2, 0, 0, 0, 0) + F(2, 0, 1, 0, 0));
    }
    if constexpr (MODE == EMF_UCT_CONTACT) {
      // upwind sign of the mass flux at the d-face of a cell (ct_emf.c:136-141)
      auto S = [&](int d, int i, int j, int k) -> int {
        const double f = sw.d[d].dv[0][at(i, j, k)];
        return f > 1.e-6 ? 1 : (f < -1.e-6 ? -1 : 0);
      };
      // face EMF minus cell EMF (the reference's DE*_D*P), and cell EMF
      // minus the face EMF one cell below along (di, dj, dk) (DE*_D*M)
      auto DP = [&](int d, int a, int c, int i, int j, int k) -> double {
        return F(d, a, i, j, k) - E(c, i, j, k);
      };
      auto DM = [&](int d, int a, int c, int i, int j, int k, int di, int dj, int dk) -> double {
        return E(c, i, j, k) - F(d, a, i - di, j - dj, k - dk);
      };
      int s, u;
      s = S(0, 0, 0, 0); u = s > 0 ? 0 : 1;
      ez += s == 0 ? 0.5 * (DP(1, 0, 2, 0, 0, 0) + DP(1, 0, 2, 1, 0, 0)) : DP(1, 0, 2, u, 0, 0);
      s = S(1, 0, 0, 0); u = s > 0 ? 0 : 1;
      ez += s == 0 ? 0.5 * (DP(0, 0, 2, 0, 0, 0) + DP(0, 0, 2, 0, 1, 0)) : DP(0, 0, 2, 0, u, 0);
      s = S(1, 1, 0, 0); u = s > 0 ? 0 : 1;
      ez -= s == 0 ? 0.5 * (DM(0, 0, 2, 1, 0, 0, 1, 0, 0) + DM(0, 0, 2, 1, 1, 0, 1, 0, 0))
                   : DM(0, 0, 2, 1, u, 0, 1, 0, 0);
      s = S(0, 0, 1, 0); u = s > 0 ? 0 : 1;
      ez -= s == 0 ? 0.5 * (DM(1, 0, 2, 0, 1, 0, 0, 1, 0) + DM(1, 0, 2, 1, 1, 0, 0, 1, 0))
                   : DM(1, 0, 2, u, 1, 0, 0, 1, 0);
      ez *= 0.25;
      if (three_d) {
        s = S(1, 0, 0, 0); u = s > 0 ? 0 : 1;
        ex += s == 0 ? 0.5 * (DP(2, 1, 0, 0, 0, 0) + DP(2, 1, 0, 0, 1, 0)) : DP(2, 1, 0, 0, u, 0);
        s = S(2, 0, 0, 0); u = s > 0 ? 0 : 1;
        ex += s == 0 ? 0.5 * (DP(1, 1, 0, 0, 0, 0) + DP(1, 1, 0, 0, 0, 1)) : DP(1, 1, 0, 0, 0, u);
        s = S(2, 0, 1, 0); u = s > 0 ? 0 : 1;
        ex -= s == 0 ? 0.5 * (DM(1, 1, 0, 0, 1, 0, 0, 1, 0) + DM(1, 1, 0, 0, 1, 1, 0, 1, 0))
                     : DM(1, 1, 0, 0, 1, u, 0, 1, 0);
        s = S(1, 0, 0, 1); u = s > 0 ? 0 : 1;
        ex -= s == 0 ? 0.5 * (DM(2, 1, 0, 0, 0, 1, 0, 0, 1) + DM(2, 1, 0, 0, 1, 1, 0, 0, 1))
                     : DM(2, 1, 0, 0, u, 1, 0, 0, 1);
        ex *= 0.25;
        s = S(0, 0, 0, 0); u = s > 0 ? 0 : 1;
        ey += s == 0 ? 0.5 * (DP(2, 0, 1, 0, 0, 0) + DP(2, 0, 1, 1, 0, 0)) : DP(2, 0, 1, u, 0, 0);
        s = S(2, 0, 0, 0); u = s > 0 ? 0 : 1;
        ey += s == 0 ? 0.5 * (DP(0, 1, 1, 0, 0, 0) + DP(0, 1, 1, 0, 0, 1)) : DP(0, 1, 1, 0, 0, u);
        s = S(2, 1, 0, 0); u = s > 0 ? 0 : 1;
        ey -= s == 0 ? 0.5 * (DM(0, 1, 1, 1, 0, 0, 1, 0, 0) + DM(0, 1, 1, 1, 0, 1, 1, 0, 0))
                     : DM(0, 1, 1, 1, 0, u, 1, 0, 0);
        s = S(0, 0, 0, 1); u = s > 0 ? 0 : 1;
        ey -= s == 0 ? 0.5 * (DM(2, 0, 1, 0, 0, 1, 0, 0, 1) + DM(2, 0, 1, 1, 0, 1, 0, 0, 1))
                     : DM(2, 0, 1, u, 0, 1, 0, 0, 1);
        ey *= 0.25;
      }
    }
  }
  e.ez[p.c] = ez;
  if (!three_d) return;
  e.ex[p.c] = ex;
  e.ey[p.c] = ey;
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
                                                  const Prims& Vout, bool* nan = nullptr)
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
  if (nan) *nan = has_nan(v);
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
