// Shared declarations between the stage kernels and the host library.
#pragma once
#include <hip/hip_runtime.h>

namespace pgpu {

constexpr int DT_HISTORY = 4096;

// Box geometry and constants passed by value to every kernel.
struct Geo {
  int nx, ny, nz, dims;
  long n;
  double dx[3];
  double inv_dx[3];   // 1.0/dx, as Grid.inv_dx (set_geometry.c:244)
  double gamma, gmm1; // g_gamma and g_gamma - 1.0
  int solver = 0;     // Riemann solver of the sweeps (SOLVER_HLLD, _HLL, _HLLC, _TVDLF, _ROE)
  int emf_avg = 4;    // EMF average (EMF_*: the CT_EMF_AVERAGE labels of ct.h:18-21)
  // A ghosted box (non-periodic boundaries): the kernels run over the whole
  // box as over a periodic one, the ghost zones are refilled between stages,
  // and the reductions keep to PLUTO's ranges: cells lo..hi are the domain.
  int ghost = 0;
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  // A stack of periodic patches (the batch): pz > 0 splits z into nz/pz
  // patches of pz planes, each periodic in z on its own, with its own
  // scalars at sc + p * sc_stride bytes (only the fields before the
  // histories).  pz = 0: one box.
  int pz = 0;
  long sc_stride = 0;
};

struct DevScalars;

// planes of one patch (the whole box when g.pz = 0)
__host__ __device__ inline int patch_depth(const Geo& g) { return g.pz > 0 ? g.pz : g.nz; }

// workgroups of a z march with kz planes per workgroup, segments not
// crossing patches
inline int zblocks(const Geo& g, int kz)
{
  const int pz = patch_depth(g);
  return (g.nz / pz) * ((pz + kz - 1) / kz);
}

// the z segment of workgroup bz: planes [k0, k1) of the patch starting at kb
struct ZSeg {
  int kb, k0, k1;
};
__device__ __forceinline__ ZSeg zseg(const Geo& g, int bz, int kz)
{
  const int pz = patch_depth(g);
  const int per = (pz + kz - 1) / kz;
  const int p = bz / per;
  ZSeg z;
  z.kb = p * pz;
  z.k0 = z.kb + (bz - p * per) * kz;
  z.k1 = min(z.k0 + kz, z.kb + pz);
  return z;
}

// plane k, periodic within the patch that starts at kb
__device__ __forceinline__ int zwrap(const Geo& g, int kb, int k)
{
  const int pz = patch_depth(g);
  int q = (k - kb) % pz;
  return kb + (q < 0 ? q + pz : q);
}

// the scalars of the patch holding plane k
__device__ __forceinline__ DevScalars* scal(DevScalars* sc, const Geo& g, int k)
{
  if (g.pz <= 0) return sc;
  return reinterpret_cast<DevScalars*>(reinterpret_cast<char*>(sc) + (long)(k / g.pz) * g.sc_stride);
}

// Cell (i, j, k) in the domain (DOM_LOOP: the CFL maximum, update_stage.c:
// 308-313; ConsToPrim3D, rk_step.c:136,189).  Always true for a periodic box.
__host__ __device__ inline bool in_dom(const Geo& g, int i, int j, int k)
{
  return !g.ghost || (i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1] &&
                      k >= g.lo[2] && k <= g.hi[2]);
}

// Interface f+1/2 along axis d (p[d] = f, unwrapped; the other two entries
// are cells) among those whose Riemann problems PLUTO solves: f in
// [beg-1, end] on the rows of the transverse ranges widened by one
// (update_stage.c:137-148,193-194), which set g_maxMach.
__host__ __device__ inline bool in_sweep(const Geo& g, int d, int i, int j, int k)
{
  if (!g.ghost) return true;
  const int p[3] = {i, j, k};
  for (int a = 0; a < 3; a++) {
    const int lo = g.lo[a] - 1, hi = a == d ? g.hi[a] : g.hi[a] + 1;
    if (p[a] < lo || p[a] > hi) return false;
  }
  return true;
}

constexpr int EMF_ARITHMETIC = 1;
constexpr int EMF_UCT0 = 2;
constexpr int EMF_UCT_CONTACT = 3;
constexpr int EMF_UCT_HLL = 4;

// Scalars that live in device memory so steps chain without host syncs.
struct DevScalars {
  double dt;                        // dt of the step being taken
  unsigned long long invdt_max;     // running max of C_dt (bit pattern, >= 0)
  unsigned long long mach_max;      // running max of the Mach estimate
  double cdt_raw_last;              // max C_dt of the last step (before /dims)
  double invdt_last;                // invDt_hyp of the last step
  double mach_last;                 // max Mach of the last step
  double dt_last;                   // dt used by the last step
  double time;
  long long step;
  unsigned int fail_cells;          // ConsToPrim floor events (cumulative)
  unsigned int nan_cells;           // stage-input cells holding a NaN (cumulative; CheckNaN)
  unsigned int nan_next;            // NaN cells of the last stage-2 output, the next
                                    // step's input: added to nan_cells when it starts
  unsigned int dt_small;            // NextTimeStep's stop (main.c:534-543) was reached
  double first_dt;                  // the run's first dt (ini->first_dt)
  double dt_small_value;            // the dt that was too small
  double dt_hist[DT_HISTORY];
  double invdt_hist[DT_HISTORY];
  double mach_hist[DT_HISTORY];
};

struct StageIn {
  const double* V[8];   // primitive state feeding the stencils (cell-centred B included)
  const double* b[3];   // staggered faces of that state
};

struct SweepOut {
  double* F[5];   // flux of rho, mx, my, mz, E at the q+1/2 interface of each cell
  double* P;      // total pressure there
  double* SL;     // max(0, -SL)
  double* SR;     // max(0,  SR)
  double* cm;     // max(|SL|, |SR|)
  double* dv[3];  // PLM velocity slopes (vp - vm) of the cell along the sweep
};

struct Sweeps {
  SweepOut d[3];
};

struct Cons {
  double* u[5];   // rho, mx, my, mz, E
};

struct Prims {
  double* v[8];
};

struct EdgeOut {
  double *ex, *ey, *ez;
};

struct FaceIn {
  const double* b[3];
};

struct FaceOut {
  double* b[3];
};

enum KernelId {
  KT_SWEEP_X = 0, KT_SWEEP_Y, KT_SWEEP_Z, KT_UPDATE, KT_EMF, KT_CT, KT_C2P, KT_STEP_END,
  KT_FLUX_XY, KT_FLUX_Z, KT_BOUNDARY, KT_COUNT
};

// Records HIP events around kernel launches (host object).
struct KernelTimer {
  virtual void begin(int id, hipStream_t s) = 0;
  virtual void end(int id, hipStream_t s) = 0;
  virtual ~KernelTimer() {}
};

int launch_stage(const Geo& g, int stage, const StageIn& in, const Sweeps& sw, const Cons& U,
                 const EdgeOut& e, const FaceIn& bin, const FaceIn& b0, const FaceOut& bout,
                 const StageIn& vn, const Prims& Vout, unsigned char* flag, DevScalars* sc,
                 hipStream_t s, KernelTimer* kt);
int launch_sweeps(const Geo& g, const StageIn& in, const Sweeps& sw, DevScalars* sc,
                  hipStream_t s, KernelTimer* kt);
int launch_update(const Geo& g, bool from_v, bool cdt, const StageIn& in, const Sweeps& sw,
                  const Cons& U, DevScalars* sc, hipStream_t s, KernelTimer* kt);
int launch_emf_ct(const Geo& g, bool blend, const StageIn& in, const Sweeps& sw, const EdgeOut& e,
                  const FaceIn& bin, const FaceIn& b0, const FaceOut& bout, DevScalars* sc,
                  hipStream_t s, KernelTimer* kt);
int launch_c2p(const Geo& g, int stage, const Cons& U, const FaceIn& b, const StageIn& vn,
               const Prims& Vout, unsigned char* flag, DevScalars* sc, hipStream_t s,
               KernelTimer* kt);
// k_emf_z tiling: workgroup bx x by (bx*by a multiple of 64, <= 256), kz
// planes per workgroup; (0, 0, 0) restores the default.  Returns -1 on bad
// values.
int set_emf_tiling(int bx, int by, int kz);
int set_emf_kernel(int kind, int kz);
// LDS-staged UCT_HLL EMF (emf_lds.hip) for a 3D box with nx even, tile shape
// 1 (64 x 4), 2 (32 x 4) or 3 (32 x 8), face data of type ft (0 double, 1
// float, 2 float with an FP32 stage input); false when the box does not
// qualify (the caller runs k_emf_z)
bool launch_emf_lds(const Geo& g, const StageIn& in, const Sweeps& sw, const EdgeOut& e,
                    hipStream_t s, int shape, int kz, KernelTimer* kt, int ft = 0);

// Branch signature of a stage (branch_sig.hip): hashes of the outcomes of
// every piecewise operation, added into *acc.
int launch_sig_sweep(const Geo& g, const StageIn& in, unsigned long long* acc, hipStream_t s);
int launch_sig_emf(const Geo& g, const StageIn& in, const Sweeps& sw, unsigned long long* acc,
                   hipStream_t s);
// A modular stage whose piecewise operations record their outcomes into
// tape (replay = 0) or take them from it (replay = 1): 4 words per cell
// (sweeps x, y, z and the EMF), HLLD + UCT_HLL, FP64, periodic.
int launch_stage_taped(const Geo& g, int stage, const StageIn& in, const Sweeps& sw,
                       const Cons& U, const EdgeOut& e, const FaceIn& bin, const FaceIn& b0,
                       const FaceOut& bout, const StageIn& vn, const Prims& Vout,
                       unsigned char* flag, DevScalars* sc, unsigned long long* tape, int replay,
                       hipStream_t s);
// CT update alone (stage 2: blended with the step's initial faces b0)
int launch_ct(const Geo& g, bool blend, const EdgeOut& e, const FaceIn& bin, const FaceIn& b0,
              const FaceOut& bout, DevScalars* sc, hipStream_t s);
void get_emf_tiling(int* bx, int* by, int* kz);
// f32_faces: the face speeds and velocity slopes in sw were written as float
// (the FP32-flux kernels), in the first half of each buffer; f32_state: the
// stage input's velocities too (the FP32-flux path's second stage).
int launch_emf(const Geo& g, const StageIn& in, const Sweeps& sw, const EdgeOut& e, hipStream_t s,
               KernelTimer* kt, bool f32_faces = false, bool f32_state = false);
// clears the flags; with sc, also moves the previous step's output NaN count
// into the stage-input count (the start of a step)
int launch_zero_flags(const Geo& g, unsigned char* flag, hipStream_t s, DevScalars* sc = nullptr);
int launch_step_end(DevScalars* sc, int dims, double cfl, double cfl_max_var, int fixed_dt,
                    hipStream_t s);

}  // namespace pgpu
