// Branch signature of a step: which side of every piecewise operation the
// step took (a diagnostic with no PLUTO counterpart).
//
// The step is a piecewise-smooth function of its input: the MC limiter, the
// min/max of the wave-speed estimates, the HLLD regions and its degeneracy
// switch, the MC_LIM2 face slopes and the max of the face speeds in UCT_HLL
// each pick one expression by a comparison.  Central finite differences of
// a many-step objective agree with the discrete adjoint (which
// differentiates the branches the forward took) only when the perturbed
// runs stay on the same pieces.  These kernels recompute every such
// comparison of a stage from the same inputs and the same expressions as
// the production kernels (mhd_device.hpp and stage_cells.hpp with a
// BranchRec attached), hash the outcomes per cell and add the hashes into
// one 64-bit word (a sum, so the order of cells does not matter).  Two runs
// with equal words took the same branches (up to hash collisions of
// probability 2^-64).
//   k_sig_sweep  every cell's own PLM limiter cases and the HLLD solve at its
//                upper face, along x, y, z, from the stage input
//   k_sig_emf    the UCT_HLL edge EMFs of every cell, from the face speeds
//                and slopes the stage's flux kernels stored
//
// Branch tape: the same outcomes, stored as bits (record) or imposed
// (replay) by the modular stage kernels' cell bodies, so that a replayed
// run evaluates the smooth piece of the step that the recorded run lay on.
// Per cell and stage four 64-bit words: one per sweep direction (the cell's
// 8 limiter cases, up to 3 bits each, then the HLLD regions and the two
// face-speed maxima of its upper face: at most 36 bits) and one for the EMF
// (12 face-speed maxima and 12 MC_LIM2 slopes: at most 48 bits).  In
// replay the slope of the next cell along the sweep comes from that cell's
// word, read in the order that cell recorded it.

#include <hip/hip_runtime.h>

#include "mhd_device.hpp"
#include "stage_cells.hpp"
#include "stage_kernels.hpp"

namespace pgpu {

namespace {

__device__ __forceinline__ unsigned long long mix64(unsigned long long x)
{
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

__device__ __forceinline__ void wave_add(unsigned long long v, unsigned long long* acc)
{
  for (int off = warpSize / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, warpSize);
  if ((threadIdx.x + threadIdx.y * blockDim.x) % warpSize == 0) atomicAdd(acc, v);
}

// cell p's limiter cases along D and the HLLD solve at its upper face, as
// sweep_cell computes them
template <int D>
__device__ __forceinline__ void sig_dir(const Geo& g, const Cell& p, const StageIn& in,
                                        BranchRec* rec)
{
  constexpr int BN = BX1 + Dir<D>::n;
  const long cm1 = nb<D>(g, p, -1), cp1 = nb<D>(g, p, 1), cp2 = nb<D>(g, p, 2);
  double vL[NV], vR[NV];
#pragma unroll
  for (int nv = 0; nv < NV; nv++) {
    const double* a = in.V[nv];
    const double v_m1 = a[cm1], v_0 = a[p.c], v_p1 = a[cp1], v_p2 = a[cp2];
    const double s0 = mc_slope(v_m1, v_0, v_p1, rec);   // this cell's cases
    const double s1 = mc_slope(v_0, v_p1, v_p2);        // the next cell's, recorded there
    vL[nv] = v_0 + s0 * 0.5;
    vR[nv] = v_p1 - s1 * 0.5;
  }
  const double bface = in.b[D][p.c];
  vL[BN] = bface;
  vR[BN] = bface;
  RiemannOut<double> r;
  riemann<SOLVER_HLLD, D, false>(vL, vR, g.gamma, g.gmm1, r, g.dims == 2, rec);
  // the face speeds the EMF averages read
  (void)max_ge(0.0, -r.SL, rec);
  (void)max_ge(0.0, r.SR, rec);
}

__global__ void __launch_bounds__(256) k_sig_sweep(Geo g, StageIn in, unsigned long long* acc)
{
  Cell p;
  unsigned long long h = 0;
  if (cell_of(g, p)) {
    BranchRec rec;
    rec.h = 0x243F6A8885A308D3ull;
    sig_dir<0>(g, p, in, &rec);
    sig_dir<1>(g, p, in, &rec);
    if (g.dims == 3) sig_dir<2>(g, p, in, &rec);
    h = mix64(rec.h + 0x9E3779B97F4A7C15ull * (unsigned long long)(p.c + 1));
  }
  wave_add(h, acc);
}

__global__ void __launch_bounds__(256) k_sig_emf(Geo g, StageIn in, Sweeps sw,
                                                 unsigned long long* acc)
{
  Cell p;
  unsigned long long h = 0;
  if (cell_of(g, p)) {
    const EmfFields f = emf_fields(in, sw);
    EmfCellLoader<double> ld{f, g, p};
    BranchRec rec;
    rec.h = 0x13198A2E03707344ull;
    double ez, ex, ey;
    emf_edges(ld, g.dims == 3, ez, ex, ey, &rec);
    h = mix64(rec.h + 0xC2B2AE3D27D4EB4Full * (unsigned long long)(p.c + 1));
  }
  wave_add(h, acc);
}

template <int D>
__global__ void __launch_bounds__(256) k_sweep_tape(Geo g, StageIn in, SweepOut out,
                                                    unsigned long long* tape, int replay)
{
  Cell p;
  if (!cell_of(g, p)) return;
  BranchRec own, next;
  own.mode = replay ? 2 : 1;
  if (replay) {
    own.word = tape[p.c];
    next.word = tape[nb<D>(g, p, 1)];
    next.mode = 2;
  }
  (void)sweep_cell<D, SOLVER_HLLD>(g, p, in, out, &own, replay ? &next : nullptr);
  if (!replay) tape[p.c] = own.word;
}

__global__ void __launch_bounds__(256) k_emf_tape(Geo g, StageIn in, Sweeps sw, EdgeOut e,
                                                  unsigned long long* tape, int replay)
{
  Cell p;
  if (!cell_of(g, p)) return;
  BranchRec rec;
  rec.mode = replay ? 2 : 1;
  if (replay) rec.word = tape[p.c];
  emf_cell<double>(g, p, in, sw, e, &rec);
  if (!replay) tape[p.c] = rec.word;
}

}  // namespace

int launch_stage_taped(const Geo& g, int stage, const StageIn& in, const Sweeps& sw,
                       const Cons& U, const EdgeOut& e, const FaceIn& bin, const FaceIn& b0,
                       const FaceOut& bout, const StageIn& vn, const Prims& Vout,
                       unsigned char* flag, DevScalars* sc, unsigned long long* tape, int replay,
                       hipStream_t s)
{
  const dim3 b(64, 4, 1), gr((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  hipLaunchKernelGGL(k_sweep_tape<0>, gr, b, 0, s, g, in, sw.d[0], tape, replay);
  hipLaunchKernelGGL(k_sweep_tape<1>, gr, b, 0, s, g, in, sw.d[1], tape + n, replay);
  if (g.dims == 3)
    hipLaunchKernelGGL(k_sweep_tape<2>, gr, b, 0, s, g, in, sw.d[2], tape + 2 * n, replay);
  int rc = (int)hipGetLastError();
  if (!rc) rc = launch_update(g, stage == 1, stage == 1, in, sw, U, sc, s, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_emf_tape, gr, b, 0, s, g, in, sw, e, tape + 3 * n, replay);
  rc = (int)hipGetLastError();
  if (!rc) rc = launch_ct(g, stage == 2, e, bin, b0, bout, sc, s);
  FaceIn bnew;
  for (int a = 0; a < 3; a++) bnew.b[a] = bout.b[a];
  if (!rc) rc = launch_c2p(g, stage, U, bnew, vn, Vout, flag, sc, s, nullptr);
  return rc;
}

int launch_sig_sweep(const Geo& g, const StageIn& in, unsigned long long* acc, hipStream_t s)
{
  const dim3 b(64, 4, 1), gr((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
  hipLaunchKernelGGL(k_sig_sweep, gr, b, 0, s, g, in, acc);
  return (int)hipGetLastError();
}

int launch_sig_emf(const Geo& g, const StageIn& in, const Sweeps& sw, unsigned long long* acc,
                   hipStream_t s)
{
  const dim3 b(64, 4, 1), gr((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
  hipLaunchKernelGGL(k_sig_emf, gr, b, 0, s, g, in, sw, acc);
  return (int)hipGetLastError();
}

}  // namespace pgpu
