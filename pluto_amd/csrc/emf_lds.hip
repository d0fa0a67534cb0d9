// UCT_HLL edge EMFs (ct_emf_average.c:179-352 with the MC_LIM2 face slopes
// of ct_stag_slopes.c:51-106), z-marching with each plane staged in LDS by
// direct global-to-LDS loads.
//
// A TX x TY workgroup owns a column tile and marches planes k0 .. k1-1.  At
// every plane the tile of the 21 EMF fields, with the in-plane halo the
// stencil reads, is copied from HBM straight into LDS with
// global_load_lds_dwordx4 (16 bytes per lane, no registers on the way), so
// each element of a plane is fetched once per workgroup, and the cells read
// their neighbours from LDS:
//
//   S fields (face speeds, velocities, velocity slopes; 18)  rows j0 .. j0+TY,
//                                                           columns i0 .. i0+TX+1
//   b_x                                                     rows j0-1 .. j0+TY+1, columns i0 .. i0+TX-1
//   b_y                                                     rows j0 .. j0+TY-1,   columns i0-2 .. i0+TX+1
//   b_z                                                     rows j0-1 .. j0+TY+1, columns i0-2 .. i0+TX+1
//
// Columns go in pairs (one 16-byte lane each; nx even keeps a pair
// contiguous across the periodic wrap).  The stencil's z offsets are taken
// as in the register kernel's split: Ex and Ey at plane k read planes k and
// k+1, so each is computed in two parts, the plane-k part while plane k is
// staged (kept in registers) and the plane-(k+1) part when plane k+1 is;
// b_x and b_y at the cell itself on planes k-1 .. k+2 (their z slopes) are
// rolled in registers.  Every quantity is the expression of emf_edges with
// the same operands in the same order, so the FP64 results are the register
// kernel's bits (the FP32-flux path's float face data are widened exactly).

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mhd_device.hpp"
#include "stage_cells.hpp"
#include "stage_kernels.hpp"

namespace pgpu {

namespace {

__device__ __forceinline__ int wrap_mod(int q, int n)
{
  const int r = q % n;
  return r < 0 ? r + n : r;
}

constexpr bool s_field(int a) { return a != F_BX && a != F_BY && a != F_BZ; }

// S fields stored as float for face-data type FT, and the slot of field a
// among the S fields of its own storage type
template <class FT>
constexpr int n_float_s()
{
  int n = 0;
  for (int a = 0; a < F_EMF_COUNT; a++) n += (s_field(a) && emf_f32<FT>(a)) ? 1 : 0;
  return n;
}

template <class FT>
constexpr int s_slot(int a)
{
  int n = 0;
  for (int b = 0; b < a; b++) n += (s_field(b) && emf_f32<FT>(b) == emf_f32<FT>(a)) ? 1 : 0;
  return n;
}

// LDS layout of one staged plane: a double region (the S fields kept as
// double, then b_x, b_y, b_z) followed by a float region (the S fields
// stored as float: face speeds and velocity slopes of the FP32-flux path,
// and its second stage's velocities)
template <int TX, int TY, class FT>
struct EmfPlane {
  static constexpr int NSF = n_float_s<FT>(), NSD = 18 - NSF;
  static constexpr int SW = TX + 2;                       // double S row: i0 .. i0+TX+1
  static constexpr int SN = (TY + 1) * SW;                // double S field: TY+1 rows
  static constexpr int FW = TX + 4;                       // float S row: i0 .. i0+TX+3
  static constexpr int FN = (TY + 1) * FW;                // float S field (floats)
  static constexpr int XW = TX, XN = (TY + 3) * XW;       // b_x: TY+3 rows
  static constexpr int YW = TX + 4, YN = TY * YW;         // b_y: TY rows, i0-2 .. i0+TX+1
  static constexpr int ZW = TX + 4, ZN = (TY + 3) * ZW;   // b_z: TY+3 rows
  static constexpr int S0 = 0, X0 = NSD * SN, Y0 = X0 + XN, Z0 = Y0 + YN, ND = Z0 + ZN;
  static constexpr int N = ND + (NSF * FN + 1) / 2;       // doubles in all
  // 16-byte lanes of each footprint and its 64-lane chunks
  static constexpr int SL = SN / 2, FL = FN / 4, XL = XN / 2, YL = YN / 2, ZL = ZN / 2;
  static constexpr int SC = (SL + 63) / 64, FC = (FL + 63) / 64, XC = (XL + 63) / 64,
                       YC = (YL + 63) / 64, ZC = (ZL + 63) / 64;
  static_assert(ND % 2 == 0 && FN % 4 == 0, "16-byte aligned regions");
};

constexpr unsigned kNoLane = 0xffffffffu;

// byte offset in a plane of lane L of a footprint of `w` elements of `es`
// bytes per row starting at (ia, ja), each lane 16 bytes; kNoLane past the
// footprint's lanes
__device__ __forceinline__ unsigned lane_offset(const Geo& g, int L, int lanes, int w, int es,
                                                int ia, int ja)
{
  if (L >= lanes) return kNoLane;
  const int per = 16 / es, pr = w / per;
  const int r = L / pr, q = L - r * pr;
  const int i = wrap_mod(ia + per * q, g.nx), j = wrap_mod(ja + r, g.ny);
  return ((unsigned)j * (unsigned)g.nx + (unsigned)i) * (unsigned)es;
}

// this thread's view of the staged plane
template <int TX, int TY, class FT>
struct EmfStaged {
  using P = EmfPlane<TX, TY, FT>;
  const double* s;
  int tx, ty;
  using R = double;   // FP32 arithmetic on float face data measured no faster
  template <int A>
  __device__ __forceinline__ R sv(int I, int J) const
  {
    if constexpr (emf_f32<FT>(A))
      return R(reinterpret_cast<const float*>(s + P::ND)[s_slot<FT>(A) * P::FN + (ty + J) * P::FW + tx + I]);
    else
      return R(s[P::S0 + s_slot<FT>(A) * P::SN + (ty + J) * P::SW + tx + I]);
  }
  __device__ __forceinline__ R bx(int J) const { return R(s[P::X0 + (ty + J + 1) * P::XW + tx]); }
  __device__ __forceinline__ R by(int I) const { return R(s[P::Y0 + ty * P::YW + tx + I + 2]); }
  __device__ __forceinline__ R bz(int I, int J) const
  {
    return R(s[P::Z0 + (ty + J + 1) * P::ZW + tx + I + 2]);
  }
};

}  // namespace

// WPE waves per SIMD: what the LDS footprint allows (64 x 4: two workgroups
// per CU, 2; 32 x 4: five, 3; 32 x 8: three, 3), which caps the VGPRs.
// One plane buffer: while a workgroup waits for its next plane, the CU
// runs the other resident workgroups.  Measured per 256^3 launch: 32 x 8
// tiles 0.84 ms, 32 x 4 0.86, 64 x 4 1.02, the register kernel 0.96; two
// buffers (the next plane in flight during the current one, one workgroup
// per CU) 1.0-1.1; 32 x 8 with XCD-contiguous tile numbering fetches 14%
// fewer bytes and runs 0.88.  FT: storage of the face data (double, float)
// and of the stage-2 input (F32State), as for k_emf_z; float fields need
// nx % 4 == 0 (four-float lanes).
template <int TX, int TY, int WPE, class FT>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE)))
k_emf_zs(Geo g, StageIn in, Sweeps sw, EdgeOut e, int kz)
{
  using P = EmfPlane<TX, TY, FT>;
  __shared__ __attribute__((aligned(16))) double s[P::N];
  constexpr int NW = TX * TY / 64;   // waves

  const int tx = threadIdx.x, ty = threadIdx.y;
  const int lane = (tx + TX * ty) & 63;
  const int wave = __builtin_amdgcn_readfirstlane((tx + TX * ty) >> 6);
  const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
  const int i = i0 + tx, j = j0 + ty;
  const bool active = i < g.nx && j < g.ny;
  const EmfFields f = emf_fields(in, sw);
  const long nxy = (long)g.nx * g.ny;
  const ZSeg zs = zseg(g, blockIdx.z, kz);   // (a patch of a stack: z periodic within it)

  // in-plane byte offsets of this lane in every chunk it may copy
  unsigned oS[P::SC], oF[P::FC], oX[P::XC], oY[P::YC], oZ[P::ZC];
#pragma unroll
  for (int c = 0; c < P::SC; c++) oS[c] = lane_offset(g, c * 64 + lane, P::SL, P::SW, 8, i0, j0);
#pragma unroll
  for (int c = 0; c < P::FC; c++) oF[c] = lane_offset(g, c * 64 + lane, P::FL, P::FW, 4, i0, j0);
#pragma unroll
  for (int c = 0; c < P::XC; c++) oX[c] = lane_offset(g, c * 64 + lane, P::XL, P::XW, 8, i0, j0 - 1);
#pragma unroll
  for (int c = 0; c < P::YC; c++) oY[c] = lane_offset(g, c * 64 + lane, P::YL, P::YW, 8, i0 - 2, j0);
#pragma unroll
  for (int c = 0; c < P::ZC; c++) oZ[c] = lane_offset(g, c * 64 + lane, P::ZL, P::ZW, 8, i0 - 2, j0 - 1);

  // copy plane k into LDS: chunk n of the list goes to wave n % NW
  auto stage = [&](int k) {
    const unsigned pl = (unsigned)zwrap(g, zs.kb, k) * (unsigned)nxy;   // plane element offset
    int n = 0;
    // dstb: byte offset of the chunk in the plane buffer
    auto chunk = [&](const void* a, unsigned off, unsigned pb, int dstb) {
      if (n++ % NW != wave) return;
      if (off != kNoLane)
        __builtin_amdgcn_global_load_lds((const void*)(reinterpret_cast<const char*>(a) + (pb + off)),
                                         (__attribute__((address_space(3))) void*)(reinterpret_cast<char*>(s) + dstb),
                                         16, 0, 0);
    };
#pragma unroll
    for (int fi = 0; fi < F_EMF_COUNT; fi++) {
      if (!s_field(fi)) continue;
      if (emf_f32<FT>(fi)) {
#pragma unroll
        for (int c = 0; c < P::FC; c++)
          chunk(f.a[fi], oF[c], pl * 4u, P::ND * 8 + (s_slot<FT>(fi) * P::FN + c * 256) * 4);
      } else {
#pragma unroll
        for (int c = 0; c < P::SC; c++)
          chunk(f.a[fi], oS[c], pl * 8u, (P::S0 + s_slot<FT>(fi) * P::SN + c * 128) * 8);
      }
    }
#pragma unroll
    for (int c = 0; c < P::XC; c++) chunk(f.a[F_BX], oX[c], pl * 8u, (P::X0 + c * 128) * 8);
#pragma unroll
    for (int c = 0; c < P::YC; c++) chunk(f.a[F_BY], oY[c], pl * 8u, (P::Y0 + c * 128) * 8);
#pragma unroll
    for (int c = 0; c < P::ZC; c++) chunk(f.a[F_BZ], oZ[c], pl * 8u, (P::Z0 + c * 128) * 8);
  };

  // LDS lookups of this thread's cell (I, J: in-plane offsets)
  const EmfStaged<TX, TY, FT> L{s, tx, ty};
  using R = typename EmfStaged<TX, TY, FT>::R;
#define S(A_, I_, J_) L.template sv<A_>(I_, J_)
#define BX(J_) L.bx(J_)
#define BY(I_) L.by(I_)
#define BZ(I_, J_) L.bz(I_, J_)

  const long cell = active ? ((long)j * g.nx + i) : 0;
  auto col = [&](const double* a, int k) -> R { return R(a[(long)zwrap(g, zs.kb, k) * nxy + cell]); };
  const int k0 = zs.k0, k1 = zs.k1;
  // b_x, b_y at the cell on planes p-1, p, p+1
  R bx_m = col(f.a[F_BX], k0 - 1), bx_0 = col(f.a[F_BX], k0), bx_p = col(f.a[F_BX], k0 + 1);
  R by_m = col(f.a[F_BY], k0 - 1), by_0 = col(f.a[F_BY], k0), by_p = col(f.a[F_BY], k0 + 1);
  // the plane-k parts of Ex(k), Ey(k)
  R x_ysr0 = 0, x_ysl0 = 0, x_ayp = 0, x_aym = 0, x_bS = 0, x_bW = 0, x_bE = 0, x_eSW = 0,
         x_eSE = 0;
  R y_axp = 0, y_axm = 0, y_xsr0 = 0, y_xsl0 = 0, y_bS = 0, y_bN = 0, y_bW = 0, y_eSW = 0,
         y_eNW = 0;

  stage(k0);
  __syncthreads();   // waits for this thread's copies, then for every wave's
  for (int p = k0; p <= k1; p++) {
    const R sBX = mc_lim2(bx_p - bx_0, bx_0 - bx_m);   // z slopes at plane p
    const R sBY = mc_lim2(by_p - by_0, by_0 - by_m);
    if (p > k0) {
      // Ex(p-1), Ey(p-1): the plane-p part (offset K = 1 of emf_edges)
      {
        const R a_xp = max_ge(x_ysr0, S(F_YSR, 0, 0));
        const R a_xm = max_ge(x_ysl0, S(F_YSL, 0, 0));
        const R bN = by_0 - R(0.5) * sBY;
        const R eNE = (S(F_VZ, 0, 1) - R(0.5) * (S(F_YDV2, 0, 1) + S(F_ZDV2, 0, 1))) * bN
                         - (S(F_VY, 0, 1) - R(0.5) * (S(F_YDV1, 0, 1) + S(F_ZDV1, 0, 1))) * x_bE;
        const R eNW = (S(F_VZ, 0, 0) + R(0.5) * (S(F_YDV2, 0, 0) - S(F_ZDV2, 0, 0))) * bN
                         - (S(F_VY, 0, 0) + R(0.5) * (S(F_YDV1, 0, 0) - S(F_ZDV1, 0, 0))) * x_bW;
        const R ex = uct_hll(a_xp, a_xm, x_ayp, x_aym, x_eSW, x_eSE, eNE, eNW, x_bS, bN,
                                  x_bW, x_bE);
        if (active) e.ex[(long)(p - 1) * nxy + cell] = ex;
      }
      __builtin_amdgcn_sched_barrier(0);
      {
        const R a_yp = max_ge(y_xsr0, S(F_XSR, 0, 0));
        const R a_ym = max_ge(y_xsl0, S(F_XSL, 0, 0));
        const R bE = bx_0 - R(0.5) * sBX;
        const R eSE = (S(F_VX, 0, 0) - R(0.5) * (S(F_ZDV0, 0, 0) - S(F_XDV0, 0, 0))) * y_bS
                         - (S(F_VZ, 0, 0) - R(0.5) * (S(F_ZDV2, 0, 0) - S(F_XDV2, 0, 0))) * bE;
        const R eNE = (S(F_VX, 1, 0) - R(0.5) * (S(F_ZDV0, 1, 0) + S(F_XDV0, 1, 0))) * y_bN
                         - (S(F_VZ, 1, 0) - R(0.5) * (S(F_ZDV2, 1, 0) + S(F_XDV2, 1, 0))) * bE;
        const R ey = uct_hll(y_axp, y_axm, a_yp, a_ym, y_eSW, eSE, eNE, y_eNW, y_bS, y_bN,
                                  y_bW, bE);
        if (active) e.ey[(long)(p - 1) * nxy + cell] = ey;
      }
    }
    if (p == k1) break;
    __builtin_amdgcn_sched_barrier(0);
    // Ez(p)
    {
      const R a_xp = max_ge(S(F_XSR, 0, 0), S(F_XSR, 0, 1));
      const R a_xm = max_ge(S(F_XSL, 0, 0), S(F_XSL, 0, 1));
      const R a_yp = max_ge(S(F_YSR, 0, 0), S(F_YSR, 1, 0));
      const R a_ym = max_ge(S(F_YSL, 0, 0), S(F_YSL, 1, 0));
      const R bS = BX(0) + R(0.5) * mc_lim2(BX(1) - BX(0), BX(0) - BX(-1));
      const R bW = BY(0) + R(0.5) * mc_lim2(BY(1) - BY(0), BY(0) - BY(-1));
      const R bN = BX(1) - R(0.5) * mc_lim2(BX(2) - BX(1), BX(1) - BX(0));
      const R bE = BY(1) - R(0.5) * mc_lim2(BY(2) - BY(1), BY(1) - BY(0));
      const R eSW = (S(F_VY, 0, 0) + R(0.5) * (S(F_XDV1, 0, 0) + S(F_YDV1, 0, 0))) * bS
                       - (S(F_VX, 0, 0) + R(0.5) * (S(F_XDV0, 0, 0) + S(F_YDV0, 0, 0))) * bW;
      const R eSE = (S(F_VY, 1, 0) - R(0.5) * (S(F_XDV1, 1, 0) - S(F_YDV1, 1, 0))) * bS
                       - (S(F_VX, 1, 0) - R(0.5) * (S(F_XDV0, 1, 0) - S(F_YDV0, 1, 0))) * bE;
      const R eNE = (S(F_VY, 1, 1) - R(0.5) * (S(F_XDV1, 1, 1) + S(F_YDV1, 1, 1))) * bN
                       - (S(F_VX, 1, 1) - R(0.5) * (S(F_XDV0, 1, 1) + S(F_YDV0, 1, 1))) * bE;
      const R eNW = (S(F_VY, 0, 1) + R(0.5) * (S(F_XDV1, 0, 1) - S(F_YDV1, 0, 1))) * bN
                       - (S(F_VX, 0, 1) + R(0.5) * (S(F_XDV0, 0, 1) - S(F_YDV0, 0, 1))) * bW;
      const R ez = uct_hll(a_xp, a_xm, a_yp, a_ym, eSW, eSE, eNE, eNW, bS, bN, bW, bE);
      if (active) e.ez[(long)p * nxy + cell] = ez;
    }
    __builtin_amdgcn_sched_barrier(0);
    // the plane-p parts of Ex(p) and Ey(p) (offset K = 0 of emf_edges)
    x_ysr0 = S(F_YSR, 0, 0);
    x_ysl0 = S(F_YSL, 0, 0);
    x_ayp = max_ge(S(F_ZSR, 0, 0), S(F_ZSR, 0, 1));
    x_aym = max_ge(S(F_ZSL, 0, 0), S(F_ZSL, 0, 1));
    x_bS = by_0 + R(0.5) * sBY;
    x_bW = BZ(0, 0) + R(0.5) * mc_lim2(BZ(0, 1) - BZ(0, 0), BZ(0, 0) - BZ(0, -1));
    x_bE = BZ(0, 1) - R(0.5) * mc_lim2(BZ(0, 2) - BZ(0, 1), BZ(0, 1) - BZ(0, 0));
    x_eSW = (S(F_VZ, 0, 0) + R(0.5) * (S(F_YDV2, 0, 0) + S(F_ZDV2, 0, 0))) * x_bS
          - (S(F_VY, 0, 0) + R(0.5) * (S(F_YDV1, 0, 0) + S(F_ZDV1, 0, 0))) * x_bW;
    x_eSE = (S(F_VZ, 0, 1) - R(0.5) * (S(F_YDV2, 0, 1) - S(F_ZDV2, 0, 1))) * x_bS
          - (S(F_VY, 0, 1) - R(0.5) * (S(F_YDV1, 0, 1) - S(F_ZDV1, 0, 1))) * x_bE;
    y_axp = max_ge(S(F_ZSR, 0, 0), S(F_ZSR, 1, 0));
    y_axm = max_ge(S(F_ZSL, 0, 0), S(F_ZSL, 1, 0));
    y_xsr0 = S(F_XSR, 0, 0);
    y_xsl0 = S(F_XSL, 0, 0);
    y_bS = BZ(0, 0) + R(0.5) * mc_lim2(BZ(1, 0) - BZ(0, 0), BZ(0, 0) - BZ(-1, 0));
    y_bW = bx_0 + R(0.5) * sBX;
    y_bN = BZ(1, 0) - R(0.5) * mc_lim2(BZ(2, 0) - BZ(1, 0), BZ(1, 0) - BZ(0, 0));
    y_eSW = (S(F_VX, 0, 0) + R(0.5) * (S(F_ZDV0, 0, 0) + S(F_XDV0, 0, 0))) * y_bS
          - (S(F_VZ, 0, 0) + R(0.5) * (S(F_ZDV2, 0, 0) + S(F_XDV2, 0, 0))) * y_bW;
    y_eNW = (S(F_VX, 1, 0) + R(0.5) * (S(F_ZDV0, 1, 0) - S(F_XDV0, 1, 0))) * y_bN
          - (S(F_VZ, 1, 0) + R(0.5) * (S(F_ZDV2, 1, 0) - S(F_XDV2, 1, 0))) * y_bW;
    // plane p+1 into LDS once every wave has read plane p; b_x, b_y of
    // plane p+2 at the cell
    const R bx_n = col(f.a[F_BX], p + 2), by_n = col(f.a[F_BY], p + 2);
    __syncthreads();
    stage(p + 1);
    __syncthreads();
    bx_m = bx_0;
    bx_0 = bx_p;
    bx_p = bx_n;
    by_m = by_0;
    by_0 = by_p;
    by_p = by_n;
  }
#undef S
#undef BX
#undef BY
#undef BZ
}

// The LDS kernel for a 3D box with nx even (column pairs; nx % 4 == 0 for
// float face data) and fields under 4 GiB (32-bit byte offsets); ft: face
// data double (0), float (1) or float with an FP32 stage input (2).  False:
// the caller takes the register kernel.
bool launch_emf_lds(const Geo& g, const StageIn& in, const Sweeps& sw, const EdgeOut& e,
                    hipStream_t s, int shape, int kz, KernelTimer* kt, int ft)
{
  if (g.dims != 3 || (g.nx & 1) || (unsigned long long)g.n * 8ull >= (1ull << 32)) return false;
  if (ft && (g.nx & 3)) return false;   // four-float lanes
  if (ft) shape = 3;                     // FP32 face data: 32 x 8 tiles only
  const int TX = shape == 1 ? 64 : 32, TY = shape == 3 ? 8 : 4;
  const long xy = (long)((g.nx + TX - 1) / TX) * ((g.ny + TY - 1) / TY);
  // fewer planes per workgroup on small boxes: keep >= 2048 workgroups
  while (kz > 1 && xy * zblocks(g, kz) < 2048) kz /= 2;
  const dim3 b(TX, TY, 1), gr((g.nx + TX - 1) / TX, (g.ny + TY - 1) / TY, zblocks(g, kz));
  if (kt) kt->begin(KT_EMF, s);
  if (ft == 1) hipLaunchKernelGGL((k_emf_zs<32, 8, 3, float>), gr, b, 0, s, g, in, sw, e, kz);
  else if (ft == 2) hipLaunchKernelGGL((k_emf_zs<32, 8, 3, F32State>), gr, b, 0, s, g, in, sw, e, kz);
  else if (shape == 1) hipLaunchKernelGGL((k_emf_zs<64, 4, 2, double>), gr, b, 0, s, g, in, sw, e, kz);
  else if (shape == 2) hipLaunchKernelGGL((k_emf_zs<32, 4, 3, double>), gr, b, 0, s, g, in, sw, e, kz);
  else hipLaunchKernelGGL((k_emf_zs<32, 8, 3, double>), gr, b, 0, s, g, in, sw, e, kz);
  if (kt) kt->end(KT_EMF, s);
  return true;
}

}  // namespace pgpu
