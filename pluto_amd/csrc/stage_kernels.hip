// Stage kernels of the unsplit RK2 MHD step (FP64 path), gfx950.
//
// One RK2 step (Time_Stepping/rk_step.c:27-254) is two stages, each:
//   sweep<x>, sweep<y>, sweep<z>   PLM + HLLD at every interface along a
//                                  direction (update_stage.c:155-242): flux,
//                                  pressure, UCT_HLL face data, velocity
//                                  slopes, Mach estimate
//   update<S>                      U += -dt/dx (F_{i+1/2} - F_{i-1/2}) and the
//                                  separate pressure-gradient term, in x, y, z
//                                  order (rhs.c:193-201); stage 1 also C_dt
//                                  and its maximum (update_stage.c:229-235,
//                                  308-313)
//   emf                            UCT_HLL edge EMFs (ct_emf_average.c:179-352)
//                                  with MC-limited staggered slopes
//                                  (ct_stag_slopes.c:51-106)
//   ct<S>                          Stokes update of the faces (ct_update.c:29-220);
//                                  stage 2 blends with the step-start faces
//                                  (rk_step.c:172-174)
//   c2p<S>                         face average of B (ct_field_average.c:64-66),
//                                  stage-2 blend of U (rk_step.c:168-170),
//                                  ConsToPrim with floors (mappers.c:88-254)
// followed by step_end, which forms invDt_hyp and the next dt on the device
// (NextTimeStep, main.c:459-532) so a run of steps never waits on the host.
//
// Layout: periodic box, no ghost zones, structure of arrays.  Cell
// c = (k*ny + j)*nx + i.  Face arrays hold the face at i+1/2 (x), j+1/2 (y)
// and k+1/2 (z) of cell c; edge arrays hold ez at (i+1/2, j+1/2, k), ex at
// (i, j+1/2, k+1/2), ey at (i+1/2, j, k+1/2).  Periodic neighbours are
// index wraps; the reference's periodic ghost zones hold exactly these
// values, so results agree bit for bit.

#include <hip/hip_runtime.h>
#include "mhd_device.hpp"
#include "stage_kernels.hpp"
#include "stage_cells.hpp"

namespace pgpu {

__device__ __forceinline__ unsigned long long as_bits(double x)
{
  return (unsigned long long)__double_as_longlong(x);
}

// Block-wide max of a non-negative double, one atomic per block.
__device__ __forceinline__ void block_max_to(double v, unsigned long long* dst)
{
  __shared__ double red[16];
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
    atomicMax(dst, as_bits(m));
  }
}

// ---------------------------------------------------------------------
// sweep: PLM + HLLD at interface (q+1/2) of every cell along direction D
// ---------------------------------------------------------------------

template <int D, int S>
__global__ void __launch_bounds__(256) k_sweep(Geo g, StageIn in, SweepOut out, DevScalars* sc)
{
  Cell p;
  double mach = 0.0;
  if (cell_of(g, p)) {
    mach = sweep_cell<D, S>(g, p, in, out);
    if (!in_sweep(g, D, p.i, p.j, p.k)) mach = 0.0;
  }
  block_max_to(mach, &sc->mach_max);
}

// ---------------------------------------------------------------------
// update: conservative accumulation in direction order, C_dt (stage 1)
// ---------------------------------------------------------------------

// FROM_V / CDT as update_cell; CDT also reduces the C_dt maximum.
template <bool FROM_V, bool CDT>
__global__ void __launch_bounds__(256) k_update(Geo g, StageIn in, Sweeps sw, Cons U, DevScalars* sc)
{
  Cell p;
  double cdt = 0.0;
  if (cell_of(g, p)) {
    cdt = update_cell<FROM_V, CDT>(g, p, in, sw, U, sc->dt);
    if (!in_dom(g, p.i, p.j, p.k)) cdt = 0.0;
  }
  if (CDT) block_max_to(cdt, &sc->invdt_max);
}

// ---------------------------------------------------------------------
// emf: UCT_HLL edge EMFs
// ---------------------------------------------------------------------

template <class FT>
__global__ void __launch_bounds__(256) k_emf(Geo g, StageIn in, Sweeps sw, EdgeOut e)
{
  Cell p;
  if (!cell_of(g, p)) return;
  emf_cell<FT>(g, p, in, sw, e);
}

// ARITHMETIC / UCT0 / UCT_CONTACT edge EMFs from the face EMFs (per cell)
template <int MODE>
__global__ void __launch_bounds__(256) k_emf_faces(Geo g, StageIn in, Sweeps sw, EdgeOut e)
{
  Cell p;
  if (!cell_of(g, p)) return;
  emf_faces_cell<MODE>(g, p, in, sw, e);
}

// z-marching EMF (3D): a workgroup covers a bx x by patch of a plane and
// marches kz planes.  The stencil reads 100 values per cell; 26 of them
// (field, i-offset, j-offset) are read on both planes k and k+1, and the
// field columns bx(i, j), by(i, j) on planes k-1 .. k+2, so a thread loads
// those once, on the leading plane, and keeps them in registers for the
// following planes: 67 loads per cell.  Every field is addressed by a 32-bit
// byte offset (in-plane part fixed per thread, plane part uniform), which
// keeps the loop's address registers to a handful.  Requires 8 n < 2^32.

// (field, i-offset, j-offset) read on planes k and k+1 by emf_edges
#define PGPU_EMF_ROLL(X)                                                                       \
  X(F_XSR, 0, 0) X(F_XSL, 0, 0) X(F_YSR, 0, 0) X(F_YSL, 0, 0)                                  \
  X(F_VX, 0, 0) X(F_VX, 1, 0) X(F_VY, 0, 0) X(F_VY, 0, 1) X(F_VZ, 0, 0) X(F_VZ, 0, 1)          \
  X(F_VZ, 1, 0)                                                                                \
  X(F_XDV0, 0, 0) X(F_XDV0, 1, 0) X(F_YDV1, 0, 0) X(F_YDV1, 0, 1) X(F_YDV2, 0, 0)              \
  X(F_YDV2, 0, 1) X(F_ZDV1, 0, 0) X(F_ZDV1, 0, 1) X(F_ZDV2, 0, 0) X(F_ZDV2, 0, 1)              \
  X(F_ZDV2, 1, 0) X(F_ZDV0, 0, 0) X(F_ZDV0, 1, 0) X(F_XDV2, 0, 0) X(F_XDV2, 1, 0)

constexpr int emf_roll_slot(int a, int i, int j)
{
  int s = 0;
#define X(A_, I_, J_)                        \
  if (a == A_ && i == I_ && j == J_) return s; \
  ++s;
  PGPU_EMF_ROLL(X)
#undef X
  return -1;
}

#define X(A_, I_, J_) +1
constexpr int kEmfRoll = 0 PGPU_EMF_ROLL(X);
#undef X

template <class FT>
struct EmfZLoader {
  const EmfFields& f;
  unsigned col[4], row[4], pk[4];   // byte offsets of the offsets -1 .. 2 along i, j, k
  double cur[kEmfRoll];             // rolled values on plane k
  double nxt[kEmfRoll];             // ... on plane k+1
  double wbx[4], wby[4];            // bx(i, j), by(i, j) on planes k-1 .. k+2

  template <int A, int I, int J, int K>
  __device__ __forceinline__ double load() const
  {
    const unsigned o = col[I + 1] + row[J + 1] + pk[K + 1];
    if constexpr (emf_f32<FT>(A))
      return double(at_byte(reinterpret_cast<const float*>(f.a[A]), o >> 1));
    else
      return at_byte(f.a[A], o);
  }

  template <int A, int I, int J, int K>
  __device__ __forceinline__ double get() const
  {
    constexpr int s = emf_roll_slot(A, I, J);
    if constexpr (A == F_BX && I == 0 && J == 0) return wbx[K + 1];
    else if constexpr (A == F_BY && I == 0 && J == 0) return wby[K + 1];
    else if constexpr (s >= 0 && K == 0) return cur[s];
    else if constexpr (s >= 0 && K == 1) return nxt[s];
    else return load<A, I, J, K>();
  }
};

// 2 waves per SIMD, the three edges scheduled one after the other (202 VGPRs,
// no spills): left alone the scheduler issues every load of an iteration at
// once and needs more than 256 VGPRs.  FT: storage type of the face speeds
// and velocity slopes.
template <class FT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_emf_z(Geo g, StageIn in, Sweeps sw, EdgeOut e, int kz)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const EmfFields f = emf_fields(in, sw);
  EmfZLoader<FT> ld{f};
  const unsigned rowb = (unsigned)g.nx * 8u;
  const unsigned planeb = rowb * (unsigned)g.ny;
#pragma unroll
  for (int o = -1; o <= 2; o++) {
    ld.col[o + 1] = (unsigned)wrap2(i, o, g.nx) * 8u;
    ld.row[o + 1] = (unsigned)wrap2(j, o, g.ny) * rowb;
  }
  const int k0 = blockIdx.z * kz;
  const int k1 = min(k0 + kz, g.nz);
  auto planes = [&](int k) {
#pragma unroll
    for (int o = -1; o <= 2; o++) ld.pk[o + 1] = (unsigned)wrap2(k, o, g.nz) * planeb;
  };
  // leading values for the first plane
  planes(k0);
#define X(A_, I_, J_) ld.cur[emf_roll_slot(A_, I_, J_)] = ld.template load<A_, I_, J_, 0>();
  PGPU_EMF_ROLL(X)
#undef X
  ld.wbx[0] = ld.template load<F_BX, 0, 0, -1>();
  ld.wbx[1] = ld.template load<F_BX, 0, 0, 0>();
  ld.wbx[2] = ld.template load<F_BX, 0, 0, 1>();
  ld.wby[0] = ld.template load<F_BY, 0, 0, -1>();
  ld.wby[1] = ld.template load<F_BY, 0, 0, 0>();
  ld.wby[2] = ld.template load<F_BY, 0, 0, 1>();
  for (int k = k0; k < k1; k++) {
    planes(k);
#define X(A_, I_, J_) ld.nxt[emf_roll_slot(A_, I_, J_)] = ld.template load<A_, I_, J_, 1>();
    PGPU_EMF_ROLL(X)
#undef X
    ld.wbx[3] = ld.template load<F_BX, 0, 0, 2>();
    ld.wby[3] = ld.template load<F_BY, 0, 0, 2>();
    double ez, ex, ey;
    emf_edges<true>(ld, true, ez, ex, ey);
    const unsigned o = ld.col[1] + ld.row[1] + ld.pk[1];
    at_byte(e.ez, o) = ez;
    at_byte(e.ex, o) = ex;
    at_byte(e.ey, o) = ey;
#pragma unroll
    for (int q = 0; q < kEmfRoll; q++) ld.cur[q] = ld.nxt[q];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      ld.wbx[q] = ld.wbx[q + 1];
      ld.wby[q] = ld.wby[q + 1];
    }
  }
}

// ---------------------------------------------------------------------
// ct: Stokes update of the faces owned by each cell
// ---------------------------------------------------------------------

template <int STAGE>
__global__ void __launch_bounds__(256) k_ct(Geo g, EdgeOut e, FaceIn bin, FaceIn b0, FaceOut bout,
                                            DevScalars* sc)
{
  Cell p;
  if (!cell_of(g, p)) return;
  ct_cell<STAGE>(g, p, e, bin, b0, bout, sc->dt);
}

// ---------------------------------------------------------------------
// c2p: face average, stage-2 blend, ConsToPrim
// ---------------------------------------------------------------------

template <int STAGE>
__global__ void __launch_bounds__(256) k_c2p(Geo g, Cons U, FaceIn b, StageIn vn,
                                             Prims Vout, unsigned char* flag,
                                             DevScalars* sc)
{
  Cell p;
  if (!cell_of(g, p)) return;
  bool nan = false;
  const unsigned char f = c2p_cell<STAGE>(g, p, U, b, vn, Vout, &nan);
  if (f) {
    flag[p.c] |= f;
    atomicAdd(&sc->fail_cells, 1u);
  }
  if (nan) atomicAdd(STAGE == 1 ? &sc->nan_cells : &sc->nan_next, 1u);
}

// Start of a step: clear the flags; the NaN cells of the previous step's
// output become stage-input NaNs now that a stage takes them as input.
__global__ void k_zero_flags(Geo g, unsigned char* flag, DevScalars* sc)
{
  if (sc && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 &&
      threadIdx.y == 0) {
    sc->nan_cells += sc->nan_next;
    sc->nan_next = 0;
  }
  Cell p;
  if (!cell_of(g, p)) return;
  flag[p.c] = 0;
}

// One thread: close the step.  invDt_hyp = max(C_dt)/DIMENSIONS
// (update_stage.c:310-311); dt_next per NextTimeStep (main.c:462-532).
__global__ void k_step_end(DevScalars* sc, int dims, double cfl, double cfl_max_var, int fixed_dt)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double inv = __longlong_as_double((long long)sc->invdt_max);
  inv /= (double)dims;
  const double mach = __longlong_as_double((long long)sc->mach_max);
  const double dt = sc->dt;
  sc->cdt_raw_last = __longlong_as_double((long long)sc->invdt_max);
  sc->invdt_last = inv;
  sc->mach_last = mach;
  sc->dt_last = dt;
  sc->time += dt;
  sc->step += 1;
  if (!fixed_dt) {
    double dt_hyp = 1.0 / inv;
    dt_hyp *= cfl;
    double dtnext = dt_hyp;
    const double cap = cfl_max_var * dt;
    dtnext = dtnext <= cap ? dtnext : cap;
    // dt too small (main.c:534-543): PLUTO quits there; record the first
    // such dt and take no more time in the steps already enqueued
    if (dtnext < sc->first_dt * 1.e-9) {
      if (!sc->dt_small) sc->dt_small_value = dtnext;
      sc->dt_small = 1;
      dtnext = 0.0;
    }
    sc->dt = dtnext;
  }
  const long long h = sc->step - 1;
  if (h >= 0 && h < DT_HISTORY) {
    sc->dt_hist[h] = dt;
    sc->invdt_hist[h] = inv;
    sc->mach_hist[h] = mach;
  }
  sc->invdt_max = 0ull;
  sc->mach_max = 0ull;
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------

static inline dim3 grid_of(const Geo& g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz); }
static const dim3 kBlock(64, 4, 1);

#define PGPU_TIMED(id, ...)              \
  do {                                   \
    if (kt) kt->begin(id, s);            \
    hipLaunchKernelGGL(__VA_ARGS__);     \
    if (kt) kt->end(id, s);              \
  } while (0)

int launch_sweeps(const Geo& g, const StageIn& in, const Sweeps& sw, DevScalars* sc,
                  hipStream_t s, KernelTimer* kt)
{
  const dim3 gr = grid_of(g);
  if (g.solver == SOLVER_HLL) {
    PGPU_TIMED(KT_SWEEP_X, (k_sweep<0, SOLVER_HLL>), gr, kBlock, 0, s, g, in, sw.d[0], sc);
    PGPU_TIMED(KT_SWEEP_Y, (k_sweep<1, SOLVER_HLL>), gr, kBlock, 0, s, g, in, sw.d[1], sc);
    if (g.dims == 3) PGPU_TIMED(KT_SWEEP_Z, (k_sweep<2, SOLVER_HLL>), gr, kBlock, 0, s, g, in, sw.d[2], sc);
  } else if (g.solver == SOLVER_HLLC) {
    PGPU_TIMED(KT_SWEEP_X, (k_sweep<0, SOLVER_HLLC>), gr, kBlock, 0, s, g, in, sw.d[0], sc);
    PGPU_TIMED(KT_SWEEP_Y, (k_sweep<1, SOLVER_HLLC>), gr, kBlock, 0, s, g, in, sw.d[1], sc);
    if (g.dims == 3) PGPU_TIMED(KT_SWEEP_Z, (k_sweep<2, SOLVER_HLLC>), gr, kBlock, 0, s, g, in, sw.d[2], sc);
  } else if (g.solver == SOLVER_TVDLF) {
    PGPU_TIMED(KT_SWEEP_X, (k_sweep<0, SOLVER_TVDLF>), gr, kBlock, 0, s, g, in, sw.d[0], sc);
    PGPU_TIMED(KT_SWEEP_Y, (k_sweep<1, SOLVER_TVDLF>), gr, kBlock, 0, s, g, in, sw.d[1], sc);
    if (g.dims == 3) PGPU_TIMED(KT_SWEEP_Z, (k_sweep<2, SOLVER_TVDLF>), gr, kBlock, 0, s, g, in, sw.d[2], sc);
  } else if (g.solver == SOLVER_ROE) {
    PGPU_TIMED(KT_SWEEP_X, (k_sweep<0, SOLVER_ROE>), gr, kBlock, 0, s, g, in, sw.d[0], sc);
    PGPU_TIMED(KT_SWEEP_Y, (k_sweep<1, SOLVER_ROE>), gr, kBlock, 0, s, g, in, sw.d[1], sc);
    if (g.dims == 3) PGPU_TIMED(KT_SWEEP_Z, (k_sweep<2, SOLVER_ROE>), gr, kBlock, 0, s, g, in, sw.d[2], sc);
  } else {
    PGPU_TIMED(KT_SWEEP_X, (k_sweep<0, SOLVER_HLLD>), gr, kBlock, 0, s, g, in, sw.d[0], sc);
    PGPU_TIMED(KT_SWEEP_Y, (k_sweep<1, SOLVER_HLLD>), gr, kBlock, 0, s, g, in, sw.d[1], sc);
    if (g.dims == 3) PGPU_TIMED(KT_SWEEP_Z, (k_sweep<2, SOLVER_HLLD>), gr, kBlock, 0, s, g, in, sw.d[2], sc);
  }
  return (int)hipGetLastError();
}

int launch_update(const Geo& g, bool from_v, bool cdt, const StageIn& in, const Sweeps& sw,
                  const Cons& U, DevScalars* sc, hipStream_t s, KernelTimer* kt)
{
  const dim3 gr = grid_of(g);
  if (from_v && cdt) PGPU_TIMED(KT_UPDATE, (k_update<true, true>), gr, kBlock, 0, s, g, in, sw, U, sc);
  else if (from_v) PGPU_TIMED(KT_UPDATE, (k_update<true, false>), gr, kBlock, 0, s, g, in, sw, U, sc);
  else if (cdt) PGPU_TIMED(KT_UPDATE, (k_update<false, true>), gr, kBlock, 0, s, g, in, sw, U, sc);
  else PGPU_TIMED(KT_UPDATE, (k_update<false, false>), gr, kBlock, 0, s, g, in, sw, U, sc);
  return (int)hipGetLastError();
}

// Tiling of k_emf_z (process-wide; pgpu_set_emf_tiling).
static int g_emf_bx = 64, g_emf_by = 4, g_emf_kz = 16;
// UCT_HLL EMF kernel of FP64 3D boxes (process-wide; pgpu_set_emf_kernel):
// 0 register-rolling k_emf_z, 1 LDS-staged k_emf_zs with 64 x 4 tiles, 2 with
// 32 x 4, 3 with 32 x 8
static int g_emf_kernel = 3, g_emf_lds_kz = 16;

int set_emf_kernel(int kind, int kz)
{
  if (kind < 0 || kind > 3 || kz < 0) return -1;
  g_emf_kernel = kind;
  g_emf_lds_kz = kz > 0 ? kz : 16;
  return 0;
}

int set_emf_tiling(int bx, int by, int kz)
{
  if (bx == 0 && by == 0 && kz == 0) {
    g_emf_bx = 64;
    g_emf_by = 4;
    g_emf_kz = 16;
    return 0;
  }
  if (bx < 1 || by < 1 || kz < 1 || bx * by > 256 || (bx * by) % 64 != 0) return -1;
  g_emf_bx = bx;
  g_emf_by = by;
  g_emf_kz = kz;
  return 0;
}

void get_emf_tiling(int* bx, int* by, int* kz)
{
  *bx = g_emf_bx;
  *by = g_emf_by;
  *kz = g_emf_kz;
}

// f32_faces: face speeds and velocity slopes stored as float (FP32 flux);
// f32_state: the velocities of the stage input too (FP32 flux, stage 2)
static bool enqueue_emf(const Geo& g, const StageIn& in, const Sweeps& sw, const EdgeOut& e,
                        hipStream_t s, KernelTimer* kt, bool f32_faces = false,
                        bool f32_state = false)
{
  if (g.pz > 0 && g.emf_avg != EMF_UCT_HLL) return false;
  if (g.emf_avg == EMF_ARITHMETIC) {
    PGPU_TIMED(KT_EMF, k_emf_faces<EMF_ARITHMETIC>, grid_of(g), kBlock, 0, s, g, in, sw, e);
  } else if (g.emf_avg == EMF_UCT0) {
    PGPU_TIMED(KT_EMF, k_emf_faces<EMF_UCT0>, grid_of(g), kBlock, 0, s, g, in, sw, e);
  } else if (g.emf_avg == EMF_UCT_CONTACT) {
    PGPU_TIMED(KT_EMF, k_emf_faces<EMF_UCT_CONTACT>, grid_of(g), kBlock, 0, s, g, in, sw, e);
  } else if ((g_emf_kernel || g.pz > 0) &&
             launch_emf_lds(g, in, sw, e, s, g_emf_kernel ? g_emf_kernel : 3, g_emf_lds_kz, kt,
                            f32_state ? 2 : (f32_faces ? 1 : 0))) {
    // LDS-staged z-march (the only EMF kernel for a stack of patches)
  } else if (g.pz > 0) {
    return false;
  } else if (g.dims == 3 && (unsigned long long)g.n * 8ull < (1ull << 32)) {
    const dim3 b(g_emf_bx, g_emf_by, 1);
    const long xy = (long)((g.nx + b.x - 1) / b.x) * ((g.ny + b.y - 1) / b.y);
    // fewer planes per workgroup on small boxes: keep >= 2048 workgroups
    // (8 per CU) in flight where the box has them
    int kz = g_emf_kz;
    while (kz > 1 && xy * ((g.nz + kz - 1) / kz) < 2048) kz /= 2;
    const dim3 gr((g.nx + b.x - 1) / b.x, (g.ny + b.y - 1) / b.y, (g.nz + kz - 1) / kz);
    if (f32_state) PGPU_TIMED(KT_EMF, k_emf_z<F32State>, gr, b, 0, s, g, in, sw, e, kz);
    else if (f32_faces) PGPU_TIMED(KT_EMF, k_emf_z<float>, gr, b, 0, s, g, in, sw, e, kz);
    else PGPU_TIMED(KT_EMF, k_emf_z<double>, gr, b, 0, s, g, in, sw, e, kz);
  } else {
    if (f32_state) PGPU_TIMED(KT_EMF, k_emf<F32State>, grid_of(g), kBlock, 0, s, g, in, sw, e);
    else if (f32_faces) PGPU_TIMED(KT_EMF, k_emf<float>, grid_of(g), kBlock, 0, s, g, in, sw, e);
    else PGPU_TIMED(KT_EMF, k_emf<double>, grid_of(g), kBlock, 0, s, g, in, sw, e);
  }
  return true;
}

int launch_emf(const Geo& g, const StageIn& in, const Sweeps& sw, const EdgeOut& e, hipStream_t s,
               KernelTimer* kt, bool f32_faces, bool f32_state)
{
  if (!enqueue_emf(g, in, sw, e, s, kt, f32_faces, f32_state)) return (int)hipErrorInvalidValue;
  return (int)hipGetLastError();
}

int launch_emf_ct(const Geo& g, bool blend, const StageIn& in, const Sweeps& sw, const EdgeOut& e,
                  const FaceIn& bin, const FaceIn& b0, const FaceOut& bout, DevScalars* sc,
                  hipStream_t s, KernelTimer* kt)
{
  // the modular sweeps store the UCT_HLL face data only
  if (g.emf_avg != EMF_UCT_HLL) return (int)hipErrorInvalidValue;
  const dim3 gr = grid_of(g);
  enqueue_emf(g, in, sw, e, s, kt);
  if (!blend) PGPU_TIMED(KT_CT, k_ct<1>, gr, kBlock, 0, s, g, e, bin, b0, bout, sc);
  else PGPU_TIMED(KT_CT, k_ct<2>, gr, kBlock, 0, s, g, e, bin, b0, bout, sc);
  return (int)hipGetLastError();
}

int launch_ct(const Geo& g, bool blend, const EdgeOut& e, const FaceIn& bin, const FaceIn& b0,
              const FaceOut& bout, DevScalars* sc, hipStream_t s)
{
  const dim3 gr = grid_of(g);
  if (!blend) hipLaunchKernelGGL(k_ct<1>, gr, kBlock, 0, s, g, e, bin, b0, bout, sc);
  else hipLaunchKernelGGL(k_ct<2>, gr, kBlock, 0, s, g, e, bin, b0, bout, sc);
  return (int)hipGetLastError();
}

int launch_c2p(const Geo& g, int stage, const Cons& U, const FaceIn& b, const StageIn& vn,
               const Prims& Vout, unsigned char* flag, DevScalars* sc, hipStream_t s,
               KernelTimer* kt)
{
  const dim3 gr = grid_of(g);
  if (stage == 1) PGPU_TIMED(KT_C2P, k_c2p<1>, gr, kBlock, 0, s, g, U, b, vn, Vout, flag, sc);
  else PGPU_TIMED(KT_C2P, k_c2p<2>, gr, kBlock, 0, s, g, U, b, vn, Vout, flag, sc);
  return (int)hipGetLastError();
}

int launch_stage(const Geo& g, int stage, const StageIn& in, const Sweeps& sw, const Cons& U,
                 const EdgeOut& e, const FaceIn& bin, const FaceIn& b0, const FaceOut& bout,
                 const StageIn& vn, const Prims& Vout, unsigned char* flag, DevScalars* sc,
                 hipStream_t s, KernelTimer* kt)
{
  int rc = launch_sweeps(g, in, sw, sc, s, kt);
  if (!rc) rc = launch_update(g, stage == 1, stage == 1, in, sw, U, sc, s, kt);
  if (!rc) rc = launch_emf_ct(g, stage == 2, in, sw, e, bin, b0, bout, sc, s, kt);
  FaceIn bnew;
  for (int d = 0; d < 3; d++) bnew.b[d] = bout.b[d];
  if (!rc) rc = launch_c2p(g, stage, U, bnew, vn, Vout, flag, sc, s, kt);
  return rc;
}
#undef PGPU_TIMED

int launch_zero_flags(const Geo& g, unsigned char* flag, hipStream_t s, DevScalars* sc)
{
  hipLaunchKernelGGL(k_zero_flags, grid_of(g), kBlock, 0, s, g, flag, sc);
  return (int)hipGetLastError();
}

int launch_step_end(DevScalars* sc, int dims, double cfl, double cfl_max_var, int fixed_dt,
                    hipStream_t s)
{
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, s, sc, dims, cfl, cfl_max_var, fixed_dt);
  return (int)hipGetLastError();
}

}  // namespace pgpu
