// CT + face average + ConsToPrim kernels of the fused stage, and the FP64
// flux launches (the flux kernels themselves are in fused_flux.hpp; the
// FP32-flux instantiations in fused_f32.hip).

#include <hip/hip_runtime.h>
#include "mhd_device.hpp"
#include "stage_cells.hpp"
#include "stage_kernels.hpp"
#include "fused_kernels.hpp"
#include "fused_flux.hpp"

namespace pgpu {

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

template <int STAGE, bool F32OUT>
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
  for (int q = 0; q < NV; q++) {
    if (F32OUT) reinterpret_cast<float*>(Vout.v[q])[c] = float(v[q]);
    else Vout.v[q][c] = v[q];
  }
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

template <int STAGE, bool F32OUT>
__global__ void __launch_bounds__(256) k_ct_c2p_z(Geo g, EdgeOut e, FaceIn bin, FaceIn b0,
                                                  FaceOut bout, Cons U, StageIn vn, Prims Vout,
                                                  unsigned char* flag, DevScalars* sc, int kz)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const ZSeg zs = zseg(g, blockIdx.z, kz);   // (a patch of a stack: z periodic within it)
  sc = scal(sc, g, zs.k0);
  const double dt = sc->dt;
  const double dtx = dt / g.dx[0], dty = dt / g.dx[1], dtz = dt / g.dx[2];
  const unsigned rowb = (unsigned)g.nx * 8u;
  const unsigned planeb = rowb * (unsigned)g.ny;
  const unsigned o00 = (unsigned)j * rowb + (unsigned)i * 8u;
  const unsigned om0 = (unsigned)j * rowb + (unsigned)wrap2(i, -1, g.nx) * 8u;   // (i-1, j)
  const unsigned o0m = (unsigned)wrap2(j, -1, g.ny) * rowb + (unsigned)i * 8u;   // (i, j-1)
  const unsigned omm = (unsigned)wrap2(j, -1, g.ny) * rowb + (unsigned)wrap2(i, -1, g.nx) * 8u;
  const int k0 = zs.k0, k1 = zs.k1;
  auto B0 = [&](int a, unsigned o) { return STAGE == 2 ? at_byte(b0.b[a], o) : 0.0; };

  // plane k0-1: edge values and the face bz
  double ey_p, eym_p, ex_p, exm_p, bz_p;
  {
    const unsigned pk = (unsigned)zwrap(g, zs.kb, k0 - 1) * planeb;
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
    for (int q = 0; q < NV; q++) {
      if (F32OUT) at_byte(reinterpret_cast<float*>(Vout.v[q]), c >> 1) = float(v[q]);
      else at_byte(Vout.v[q], c) = v[q];
    }
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
                  unsigned char* flag, DevScalars* sc, hipStream_t s, KernelTimer* kt,
                  bool f32_out)
{
  if (f32_out && stage != 1) return (int)hipErrorInvalidValue;
  const bool zmarch = g.dims == 3 && (unsigned long long)g.n * 8ull < (1ull << 32);
  if (g.pz > 0 && !zmarch) return (int)hipErrorInvalidValue;   // stacks: the z march only
  // rows of 128 cells x 2 (less y halo re-read than 64 x 4: 0.81 -> 0.78 ms
  // at 256^3) where that wastes no more lanes than 64 x 4 (261-cell rows of
  // a ghosted box: 64 x 4)
  const int bx = (g.nx + 127) / 128 * 128 <= (g.nx + 63) / 64 * 64 ? 128 : 64, by = 256 / bx;
  const dim3 b(bx, by), gr((g.nx + bx - 1) / bx, (g.ny + by - 1) / by, g.nz);
  if (kt) kt->begin(KT_CT, s);
  if (zmarch) {
    // planes per workgroup: up to 16, fewer while the box would get fewer
    // than 2048 workgroups
    const long xy = (long)gr.x * gr.y;
    int kz = 16;
    while (kz > 1 && xy * zblocks(g, kz) < 2048) kz /= 2;
    const dim3 gz(gr.x, gr.y, zblocks(g, kz));
    if (f32_out)
      hipLaunchKernelGGL((k_ct_c2p_z<1, true>), gz, b, 0, s, g, e, bin, b0, bout, U, vn, Vout, flag, sc, kz);
    else if (stage == 1)
      hipLaunchKernelGGL((k_ct_c2p_z<1, false>), gz, b, 0, s, g, e, bin, b0, bout, U, vn, Vout, flag, sc, kz);
    else
      hipLaunchKernelGGL((k_ct_c2p_z<2, false>), gz, b, 0, s, g, e, bin, b0, bout, U, vn, Vout, flag, sc, kz);
  } else if (f32_out) {
    hipLaunchKernelGGL((k_ct_c2p<1, true>), gr, b, 0, s, g, e, bin, b0, bout, U, vn, Vout, flag, sc);
  } else if (stage == 1) {
    hipLaunchKernelGGL((k_ct_c2p<1, false>), gr, b, 0, s, g, e, bin, b0, bout, U, vn, Vout, flag, sc);
  } else {
    hipLaunchKernelGGL((k_ct_c2p<2, false>), gr, b, 0, s, g, e, bin, b0, bout, U, vn, Vout, flag, sc);
  }
  if (kt) kt->end(KT_CT, s);
  return (int)hipGetLastError();
}

static int g_flux_xy_kind = 1, g_flux_xy_seg = 0;

int set_flux_xy_kernel(int kind, int seg)
{
  if (kind < 0 || kind > 1 || seg < 0) return -1;
  g_flux_xy_kind = kind;
  g_flux_xy_seg = seg;
  return 0;
}

int flux_xy_kind() { return g_flux_xy_kind; }

int flux_xy_seg(const Geo& g)
{
  if (g_flux_xy_seg > 0) return g_flux_xy_seg;
  // 32 rows, halved (down to 4) while the box would get fewer than 2048
  // waves (8 per CU); measured: 256^3 best at 32 (8192 waves), 128^3 at 16
  const long waves_per_row = (g.nx + 63) / 64;
  int seg = 32;
  while (seg > 4 && (long)((g.ny + seg - 1) / seg) * g.nz * waves_per_row < 2048) seg /= 2;
  return seg;
}

int launch_flux_fused(const Geo& g, int stage, bool fp32_flux, const StageIn& in, const Cons& U,
                      const FaceData& fd, double* cdt_xy, DevScalars* sc, hipStream_t s,
                      KernelTimer* kt)
{
  // a stack of patches runs on the row kernel (its z march handles stacks);
  // the z march takes 32-bit offsets: 3D fields under 4 GiB
  if (g.pz > 0 && (g.dims != 3 || !flux_rows_ok(g, fp32_flux ? 4 : 8, g.solver)))
    return (int)hipErrorInvalidValue;
  if (g.dims == 3 && g.n * 8 >= (1l << 32)) return (int)hipErrorInvalidValue;
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
  if (fp32_flux) return launch_flux_fused_f32(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  if (g.solver == SOLVER_HLL) launch_flux_fused_t<double, SOLVER_HLL>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  else if (g.solver == SOLVER_HLLC) launch_flux_fused_t<double, SOLVER_HLLC>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  else if (g.solver == SOLVER_TVDLF) launch_flux_fused_t<double, SOLVER_TVDLF>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  else if (g.solver == SOLVER_ROE) launch_flux_fused_t<double, SOLVER_ROE>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  else launch_flux_fused_t<double, SOLVER_HLLD>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  return (int)hipGetLastError();
}

}  // namespace pgpu
