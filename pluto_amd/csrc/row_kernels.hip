// Row-granular kernels behind PLUTO's row-level entry points
// (HLLD_Solver, HLL_Solver, HLLC_Solver, States, PrimToCons, ConsToPrim) and the face average of
// CT_AverageMagneticField.  Rows are PLUTO's AoS layout, [n][NVAR] with
// NVAR = 8 in the canonical order.

#include <hip/hip_runtime.h>
#include "mhd_device.hpp"
#include "row_kernels.hpp"
#include "stage_cells.hpp"

namespace pgpu {

template <int D, int S>
__global__ void k_riemann_rows(int n, const double* __restrict__ vl, const double* __restrict__ vr,
                            double gamma, double* flux, double* press, double* sl, double* sr,
                            double* cmax, double* mach)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a[NV], b[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    a[q] = vl[(long)i * NV + q];
    b[q] = vr[(long)i * NV + q];
  }
  RiemannOut<double> o;
  riemann<S, D, true>(a, b, gamma, gamma - 1.0, o);
#pragma unroll
  for (int q = 0; q < NV; q++) flux[(long)i * NV + q] = o.f[q];
  press[i] = o.press;
  sl[i] = o.SL;
  sr[i] = o.SR;
  cmax[i] = o.cmax;
  mach[i] = o.mach;
}

// PLM of cells 0..n-1 from v[-1..n] (v points at cell -1); bn[-1..n-1]
// holds the staggered normal field at the faces i+1/2 (bn points at face
// -1/2).  Outputs vp (face i+1/2) and vm (face i-1/2) with the normal field
// replaced as in plm_states.c:271-274, and their conservative forms.
template <int D>
__global__ void k_plm_rows(int n, const double* __restrict__ v, const double* __restrict__ bn,
                           double gamma, double* vp, double* vm, double* up, double* um)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int BN = BX1 + Dir<D>::n;
  const double* c = v + (long)(i + 1) * NV;
  double p[NV], m[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    const double s = mc_slope(c[q - NV], c[q], c[q + NV]);
    p[q] = c[q] + s * 0.5;
    m[q] = c[q] - s * 0.5;
  }
  p[BN] = bn[i + 1];
  m[BN] = bn[i];
  double u[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    vp[(long)i * NV + q] = p[q];
    vm[(long)i * NV + q] = m[q];
  }
  prim2cons(p, u, gamma - 1.0);
#pragma unroll
  for (int q = 0; q < NV; q++) up[(long)i * NV + q] = u[q];
  prim2cons(m, u, gamma - 1.0);
#pragma unroll
  for (int q = 0; q < NV; q++) um[(long)i * NV + q] = u[q];
}

__global__ void k_p2c_rows(int n, const double* __restrict__ v, double gamma, double* u)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a[NV], b[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) a[q] = v[(long)i * NV + q];
  prim2cons(a, b, gamma - 1.0);
#pragma unroll
  for (int q = 0; q < NV; q++) u[(long)i * NV + q] = b[q];
}

__global__ void k_c2p_rows(int n, double* u, double gamma, double* v, unsigned char* flag)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a[NV], b[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) a[q] = u[(long)i * NV + q];
  flag[i] = cons2prim(a, b, gamma - 1.0);
#pragma unroll
  for (int q = 0; q < NV; q++) {
    u[(long)i * NV + q] = a[q];
    v[(long)i * NV + q] = b[q];
  }
}

// Face average on a packed box: faces f[a] hold (mx+1)*(my)*(mz) etc. with
// one extra face on the low side of their own axis; out is [mz][my][mx][3].
__global__ void k_face_average_box(int mx, int my, int mz, const double* __restrict__ fx,
                                   const double* __restrict__ fy, const double* __restrict__ fz,
                                   int dims, double* out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = blockIdx.z;
  if (i >= mx) return;
  const long c = ((long)k * my + j) * mx + i;
  // fx: [mz][my][mx+1], face index i+1 is the right face of cell i
  const long cx = ((long)k * my + j) * (mx + 1) + i;
  const long cy = ((long)k * (my + 1) + j) * mx + i;
  out[c * 3 + 0] = 0.5 * (fx[cx + 1] + fx[cx]);
  out[c * 3 + 1] = 0.5 * (fy[cy + mx] + fy[cy]);
  if (dims == 3) {
    const long cz = ((long)k * my + j) * mx + i;
    out[c * 3 + 2] = 0.5 * (fz[cz + (long)mx * my] + fz[cz]);
  }
}

template <int S>
static void riemann_rows(int dir, int n, const double* vl, const double* vr, double gamma,
                         double* flux, double* press, double* sl, double* sr, double* cmax,
                         double* mach, hipStream_t s)
{
  const dim3 b(64), gr((n + 63) / 64);
  if (dir == 0) hipLaunchKernelGGL((k_riemann_rows<0, S>), gr, b, 0, s, n, vl, vr, gamma, flux, press, sl, sr, cmax, mach);
  else if (dir == 1) hipLaunchKernelGGL((k_riemann_rows<1, S>), gr, b, 0, s, n, vl, vr, gamma, flux, press, sl, sr, cmax, mach);
  else hipLaunchKernelGGL((k_riemann_rows<2, S>), gr, b, 0, s, n, vl, vr, gamma, flux, press, sl, sr, cmax, mach);
}

int launch_riemann_rows(int solver, int dir, int n, const double* vl, const double* vr,
                        double gamma, double* flux, double* press, double* sl, double* sr,
                        double* cmax, double* mach, hipStream_t s)
{
  if (solver == SOLVER_HLL) riemann_rows<SOLVER_HLL>(dir, n, vl, vr, gamma, flux, press, sl, sr, cmax, mach, s);
  else if (solver == SOLVER_HLLC) riemann_rows<SOLVER_HLLC>(dir, n, vl, vr, gamma, flux, press, sl, sr, cmax, mach, s);
  else if (solver == SOLVER_TVDLF) riemann_rows<SOLVER_TVDLF>(dir, n, vl, vr, gamma, flux, press, sl, sr, cmax, mach, s);
  else if (solver == SOLVER_ROE) riemann_rows<SOLVER_ROE>(dir, n, vl, vr, gamma, flux, press, sl, sr, cmax, mach, s);
  else riemann_rows<SOLVER_HLLD>(dir, n, vl, vr, gamma, flux, press, sl, sr, cmax, mach, s);
  return (int)hipGetLastError();
}

int launch_plm_rows(int dir, int n, const double* v, const double* bn, double gamma,
                    double* vp, double* vm, double* up, double* um, hipStream_t s)
{
  const dim3 b(64), gr((n + 63) / 64);
  if (dir == 0) hipLaunchKernelGGL(k_plm_rows<0>, gr, b, 0, s, n, v, bn, gamma, vp, vm, up, um);
  else if (dir == 1) hipLaunchKernelGGL(k_plm_rows<1>, gr, b, 0, s, n, v, bn, gamma, vp, vm, up, um);
  else hipLaunchKernelGGL(k_plm_rows<2>, gr, b, 0, s, n, v, bn, gamma, vp, vm, up, um);
  return (int)hipGetLastError();
}

int launch_p2c_rows(int n, const double* v, double gamma, double* u, hipStream_t s)
{
  hipLaunchKernelGGL(k_p2c_rows, dim3((n + 63) / 64), dim3(64), 0, s, n, v, gamma, u);
  return (int)hipGetLastError();
}

int launch_c2p_rows(int n, double* u, double gamma, double* v, unsigned char* flag, hipStream_t s)
{
  hipLaunchKernelGGL(k_c2p_rows, dim3((n + 63) / 64), dim3(64), 0, s, n, u, gamma, v, flag);
  return (int)hipGetLastError();
}

int launch_face_average_box(int mx, int my, int mz, const double* fx, const double* fy,
                            const double* fz, int dims, double* out, hipStream_t s)
{
  hipLaunchKernelGGL(k_face_average_box, dim3((mx + 63) / 64, my, mz), dim3(64), 0, s, mx, my, mz,
                     fx, fy, fz, dims, out);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------
// CT on PLUTO's ghosted boxes (the CT_ComputeEMF / CT_Update entry points)
// ---------------------------------------------------------------------

namespace {

// emf_edges loader over a box without wraps: the caller's ranges keep every
// offset in [-1, 2] inside the box
struct BoxLoader {
  const EmfFields& f;
  long c, sy, sz;
  template <int A, int I, int J, int K>
  __device__ __forceinline__ double get() const
  {
    return f.a[A][c + I + J * sy + K * sz];
  }
};

struct BoxRange {
  int n[3], lo[3], hi[3];
};

// CT_GetStagSlopes + CT_EMF_HLL_Solver (ct_stag_slopes.c:51-106,
// ct_emf_average.c:179-352) at the edges of cell (i, j, k) of the range
__global__ void __launch_bounds__(256) k_ct_emf_box(BoxRange r, EmfFields f, double* ex,
                                                    double* ey, double* ez)
{
  const int i = r.lo[0] + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = r.lo[1] + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = r.lo[2] + blockIdx.z;
  if (i > r.hi[0] || j > r.hi[1]) return;
  const long c = ((long)k * r.n[1] + j) * r.n[0] + i;
  BoxLoader ld{f, c, r.n[0], (long)r.n[0] * r.n[1]};
  double e_z, e_x, e_y;
  emf_edges(ld, true, e_z, e_x, e_y);
  ez[c] = e_z;
  ex[c] = e_x;
  ey[c] = e_y;
}

struct FacePtrs {
  const double* b0[3];
  double* b[3];
};

// CT_Update (ct_update.c:79-190, Cartesian): each face component over its
// own range inside the EMF range, with the reference's per-index widths
__global__ void __launch_bounds__(256) k_ct_update_box(BoxRange r, double dt, const double* dx1,
                                                       const double* dx2, const double* dx3,
                                                       const double* ex, const double* ey,
                                                       const double* ez, FacePtrs fp)
{
  const int i = r.lo[0] + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = r.lo[1] + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = r.lo[2] + blockIdx.z;
  if (i > r.hi[0] || j > r.hi[1]) return;
  const long sy = r.n[0], sz = (long)r.n[0] * r.n[1];
  const long c = k * sz + j * sy + i;
  if (k >= r.lo[2] + 1 && j >= r.lo[1] + 1) {
    const double rhs_x = 0.0 - dt / dx2[j] * (ez[c] - ez[c - sy]) + dt / dx3[k] * (ey[c] - ey[c - sz]);
    fp.b[0][c] = fp.b0[0][c] + rhs_x;
  }
  if (k >= r.lo[2] + 1 && i >= r.lo[0] + 1) {
    const double rhs_y = dt / dx1[i] * (ez[c] - ez[c - 1]) - dt / dx3[k] * (ex[c] - ex[c - sz]);
    fp.b[1][c] = fp.b0[1][c] + rhs_y;
  }
  if (j >= r.lo[1] + 1 && i >= r.lo[0] + 1) {
    const double rhs_z = -dt / dx1[i] * (ey[c] - ey[c - 1]) + dt / dx2[j] * (ex[c] - ex[c - sy]);
    fp.b[2][c] = fp.b0[2][c] + rhs_z;
  }
}

bool range_ok(const int n[3], const int lo[3], const int hi[3])
{
  for (int a = 0; a < 3; a++)
    if (lo[a] < 1 || hi[a] > n[a] - 3 || lo[a] > hi[a]) return false;
  return true;
}

}  // namespace

int launch_ct_emf_box(const int n[3], const int lo[3], const int hi[3], const double* const* f21,
                      double* ex, double* ey, double* ez, hipStream_t s)
{
  if (!range_ok(n, lo, hi)) return (int)hipErrorInvalidValue;
  BoxRange r;
  for (int a = 0; a < 3; a++) {
    r.n[a] = n[a];
    r.lo[a] = lo[a];
    r.hi[a] = hi[a];
  }
  EmfFields f;
  for (int q = 0; q < F_EMF_COUNT; q++) f.a[q] = f21[q];
  const dim3 b(64, 4), g((hi[0] - lo[0] + 64) / 64, (hi[1] - lo[1] + 4) / 4, hi[2] - lo[2] + 1);
  hipLaunchKernelGGL(k_ct_emf_box, g, b, 0, s, r, f, ex, ey, ez);
  return (int)hipGetLastError();
}

int launch_ct_update_box(const int n[3], const int lo[3], const int hi[3], double dt,
                         const double* dx1, const double* dx2, const double* dx3,
                         const double* ex, const double* ey, const double* ez,
                         const double* const* b0, double* const* b, hipStream_t s)
{
  if (!range_ok(n, lo, hi)) return (int)hipErrorInvalidValue;
  BoxRange r;
  for (int a = 0; a < 3; a++) {
    r.n[a] = n[a];
    r.lo[a] = lo[a];
    r.hi[a] = hi[a];
  }
  FacePtrs fp;
  for (int a = 0; a < 3; a++) {
    fp.b0[a] = b0[a];
    fp.b[a] = b[a];
  }
  const dim3 bl(64, 4), g((hi[0] - lo[0] + 64) / 64, (hi[1] - lo[1] + 4) / 4, hi[2] - lo[2] + 1);
  hipLaunchKernelGGL(k_ct_update_box, g, bl, 0, s, r, dt, dx1, dx2, dx3, ex, ey, ez, fp);
  return (int)hipGetLastError();
}

}  // namespace pgpu
