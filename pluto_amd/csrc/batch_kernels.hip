// Batched patches: many independent periodic cubes of different sizes
// advanced together, one persistent launch per RK stage.
//
// Storage: every zone array (V, V1, U, faces, sweep data, edges) holds the
// patches back to back, patch p at offset off[p] with n_p^3 zones in the
// [z][y][x] order of a single domain, so a patch is a periodic box that the
// cell functions of stage_cells.hpp handle unchanged (results are bit for
// bit those of a single domain of that size).
//
// Work, two ways, both captured once per step in a hipGraph and replayed:
//
//   stacked (mode 1, the default for patches with an even edge of 4 to 256
//   cells): the patches of one edge size lie back to back, which makes them
//   one box of n x n x (count * n) cells split in z into periodic patches
//   (Geo::pz); the fused stage kernels of a single domain (row-marching x/y
//   flux, z march, LDS EMF, z-marching CT + ConsToPrim) run over each such
//   stack with z wrapped within each patch and each patch's dt and
//   reductions in its own scalars (Geo::sc_stride).  A step is 2 x 4
//   launches per edge size plus a scalar kernel before and after.
//
//   persistent (mode 0, and patches of other edges in mode 1): a grid of
//   resident workgroups takes patches from a device counter (largest first)
//   and runs the whole stage on its patch -- sweeps, update, EMF, CT,
//   ConsToPrim -- with the modular cell bodies, separated by workgroup
//   barriers, since every phase only reads data of its own patch.
//
// Each patch keeps its own time step: stage 1 reduces the patch's C_dt
// maximum, stage 2 (or the step's last kernel) closes the step with
// NextTimeStep's rule (main.c:459-532) for that patch.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <numeric>
#include <vector>

#include "../../include/pluto_amd.h"
#include "domain.hpp"
#include "fused_kernels.hpp"
#include "mhd_device.hpp"
#include "stage_cells.hpp"
#include "stage_kernels.hpp"

namespace pgpu {

struct PatchDesc {
  int n;          // edge (cells)
  int pad;
  long off;       // zone offset in every array
  double dx;      // 1/n: each patch is a unit periodic cube
};

struct PatchScalars {
  double dt;          // dt of the step being taken
  double cdt_max;     // max C_dt of stage 1
  double invdt_last;  // invDt_hyp of the last step
  double mach_last;   // max Mach of the last step
  double dt_last;
  double time;
  long long step;
  unsigned int fail_cells;
  unsigned int dt_small;   // NextTimeStep's stop (main.c:534-543) was reached
  double first_dt;         // the patch's first dt
};

struct BatchArrays {
  double* V[8];
  double* V1[8];
  double* U[5];
  double* bA[3];
  double* bB[3];
  double* sw[3][12];   // per direction: F[5], P, SL, SR, cm, dv[3]
  double* e[3];        // ex, ey, ez
  unsigned char* flag;
};

struct BatchCtl {
  const PatchDesc* pd;
  const int* order;          // patch ids, largest first
  PatchScalars* ps;
  unsigned int* queue;       // [0] stage-1 counter, [1] stage-2 counter
  int npatch;
  int fixed_dt;
  int tx, ty, tz;            // cell tile of one pass (tx*ty*tz = workgroup size), or 0: linear
  double gamma, gmm1, cfl, cfl_max_var;
};

namespace {

__device__ __forceinline__ double block_max_d(double v, double* red)
{
  const int lane = threadIdx.x;
  const int wave = lane / warpSize;
  const int nwave = (blockDim.x + warpSize - 1) / warpSize;
  for (int off = warpSize / 2; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, warpSize);
    v = o > v ? o : v;
  }
  __syncthreads();
  if ((lane % warpSize) == 0) red[wave] = v;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < nwave; w++) m = red[w] > m ? red[w] : m;
  return m;
}

// The cells of an n^3 patch that a workgroup's threads take pass by pass.
// Linear (tx = 0): cells threadIdx.x, threadIdx.x + blockDim.x, ... in
// storage order, (i, j, k) stepped incrementally.  Tiled: each pass is one
// tx x ty x tz box of cells (tx*ty*tz = blockDim.x), the boxes in storage
// order; a box keeps the stencil neighbours of its cells close in time,
// cells of a box beyond the patch are skipped (in() false).  The order of
// the cells does not change any result: each phase writes per-cell outputs
// and reduces with max and integer sums.
struct CellIter {
  Cell p;
  int n, di, dj, dk, stride, total;
  int tx, ty, tz, ti, tj, tk, b0, b1, nb0, nb1, tile, ntile;
  __device__ __forceinline__ CellIter(int n_, int tx_, int ty_, int tz_)
      : n(n_), tx(tx_), ty(ty_), tz(tz_)
  {
    const int t = threadIdx.x;
    if (tx == 0) {
      stride = blockDim.x;
      total = n * n * n;
      p.i = t % n;
      p.j = (t / n) % n;
      p.k = t / (n * n);
      p.c = t;
      di = stride % n;
      dj = (stride / n) % n;
      dk = stride / (n * n);
      return;
    }
    ti = t % tx;
    tj = (t / tx) % ty;
    tk = t / (tx * ty);
    nb0 = (n + tx - 1) / tx;
    nb1 = (n + ty - 1) / ty;
    ntile = nb0 * nb1 * ((n + tz - 1) / tz);
    tile = 0;
    b0 = b1 = 0;
    p.k = tk;
    place();
  }
  __device__ __forceinline__ void place()
  {
    p.i = b0 * tx + ti;
    p.j = b1 * ty + tj;
    p.c = ((long)p.k * n + p.j) * n + p.i;
  }
  __device__ __forceinline__ bool ok() const { return tx == 0 ? p.c < total : tile < ntile; }
  __device__ __forceinline__ bool in() const { return tx == 0 || (p.i < n && p.j < n && p.k < n); }
  __device__ __forceinline__ void next()
  {
    if (tx != 0) {
      tile++;
      if (++b0 == nb0) {
        b0 = 0;
        if (++b1 == nb1) {
          b1 = 0;
          p.k += tz;
        }
      }
      place();
      return;
    }
    p.c += stride;
    p.i += di;
    int carry = 0;
    if (p.i >= n) {
      p.i -= n;
      carry = 1;
    }
    p.j += dj + carry;
    carry = 0;
    if (p.j >= n) {
      p.j -= n;
      carry = 1;
    }
    p.k += dk + carry;
  }
};

}  // namespace

// MAXT: the largest workgroup the instance runs.  At 1024 threads the
// compiler has 128 VGPRs and spills the sweep (about 340 bytes per lane); at
// 512 it keeps everything in about 220 VGPRs (2 waves per SIMD).  The
// launcher picks the instance by the workgroup size, the autotuner by time.
template <int STAGE, int MAXT>
__global__ void __launch_bounds__(MAXT) k_batch_stage(BatchArrays A, BatchCtl C)
{
  __shared__ int s_q;
  __shared__ double red[16];
  if (blockIdx.x == 0 && threadIdx.x == 0) C.queue[STAGE == 1 ? 1 : 0] = 0u;   // the other stage's counter
  for (;;) {
    if (threadIdx.x == 0) s_q = (int)atomicAdd(&C.queue[STAGE - 1], 1u);
    __syncthreads();
    const int q = s_q;
    __syncthreads();
    if (q >= C.npatch) break;
    const int pid = C.order[q];
    const PatchDesc d = C.pd[pid];
    Geo g;
    g.nx = g.ny = g.nz = d.n;
    g.dims = 3;
    g.n = (long)d.n * d.n * d.n;
    for (int a = 0; a < 3; a++) {
      g.dx[a] = d.dx;
      g.inv_dx[a] = 1.0 / d.dx;
    }
    g.gamma = C.gamma;
    g.gmm1 = C.gmm1;
    const double dt = C.ps[pid].dt;

    // views of this patch
    StageIn in, vn;
    Sweeps sw;
    Cons U;
    EdgeOut e;
    FaceIn bin, b0;
    FaceOut bout;
    Prims Vout;
    for (int v = 0; v < 8; v++) {
      in.V[v] = (STAGE == 1 ? A.V[v] : A.V1[v]) + d.off;
      vn.V[v] = A.V[v] + d.off;
      Vout.v[v] = (STAGE == 1 ? A.V1[v] : A.V[v]) + d.off;
    }
    for (int a = 0; a < 3; a++) {
      in.b[a] = (STAGE == 1 ? A.bA[a] : A.bB[a]) + d.off;
      vn.b[a] = A.bA[a] + d.off;
      bin.b[a] = in.b[a];
      b0.b[a] = A.bA[a] + d.off;
      bout.b[a] = (STAGE == 1 ? A.bB[a] : A.bA[a]) + d.off;
    }
    for (int dd = 0; dd < 3; dd++) {
      for (int q5 = 0; q5 < 5; q5++) sw.d[dd].F[q5] = A.sw[dd][q5] + d.off;
      sw.d[dd].P = A.sw[dd][5] + d.off;
      sw.d[dd].SL = A.sw[dd][6] + d.off;
      sw.d[dd].SR = A.sw[dd][7] + d.off;
      sw.d[dd].cm = A.sw[dd][8] + d.off;
      for (int k = 0; k < 3; k++) sw.d[dd].dv[k] = A.sw[dd][9 + k] + d.off;
    }
    for (int q5 = 0; q5 < 5; q5++) U.u[q5] = A.U[q5] + d.off;
    e.ex = A.e[0] + d.off;
    e.ey = A.e[1] + d.off;
    e.ez = A.e[2] + d.off;
    unsigned char* flag = A.flag + d.off;

    // sweeps
    double mach = 0.0;
    for (CellIter it(d.n, C.tx, C.ty, C.tz); it.ok(); it.next()) {
      if (!it.in()) continue;
      const Cell& p = it.p;
      if (STAGE == 1) flag[p.c] = 0;
      const double m0 = sweep_cell<0>(g, p, in, sw.d[0]);
      const double m1 = sweep_cell<1>(g, p, in, sw.d[1]);
      const double m2 = sweep_cell<2>(g, p, in, sw.d[2]);
      mach = m0 > mach ? m0 : mach;
      mach = m1 > mach ? m1 : mach;
      mach = m2 > mach ? m2 : mach;
    }
    __syncthreads();
    // update
    double cdt = 0.0;
    for (CellIter it(d.n, C.tx, C.ty, C.tz); it.ok(); it.next()) {
      if (!it.in()) continue;
      const Cell& p = it.p;
      const double c1 = update_cell<STAGE == 1, STAGE == 1>(g, p, in, sw, U, dt);
      cdt = c1 > cdt ? c1 : cdt;
    }
    __syncthreads();
    // EMF
    for (CellIter it(d.n, C.tx, C.ty, C.tz); it.ok(); it.next())
      if (it.in()) emf_cell(g, it.p, in, sw, e);
    __syncthreads();
    // CT
    for (CellIter it(d.n, C.tx, C.ty, C.tz); it.ok(); it.next())
      if (it.in()) ct_cell<STAGE>(g, it.p, e, bin, b0, bout, dt);
    __syncthreads();
    // face average, blend, ConsToPrim
    FaceIn bnew;
    for (int a = 0; a < 3; a++) bnew.b[a] = bout.b[a];
    unsigned int fails = 0;
    for (CellIter it(d.n, C.tx, C.ty, C.tz); it.ok(); it.next()) {
      if (!it.in()) continue;
      const unsigned char f = c2p_cell<STAGE>(g, it.p, U, bnew, vn, Vout);
      if (f) {
        flag[it.p.c] |= f;
        fails++;
      }
    }
    const double mach_b = block_max_d(mach, red);
    const double cdt_b = block_max_d(cdt, red);
    if (fails) atomicAdd(&C.ps[pid].fail_cells, fails);
    if (threadIdx.x == 0) {
      PatchScalars& s = C.ps[pid];
      if (STAGE == 1) {
        s.cdt_max = cdt_b;
        s.mach_last = mach_b;
      } else {
        // close the step (k_step_end's rule for this patch)
        const double mach_step = mach_b > s.mach_last ? mach_b : s.mach_last;
        double inv = s.cdt_max;
        inv /= 3.0;
        s.invdt_last = inv;
        s.mach_last = mach_step;
        s.dt_last = dt;
        s.time += dt;
        s.step += 1;
        if (!C.fixed_dt) {
          double dt_hyp = 1.0 / inv;
          dt_hyp *= C.cfl;
          double dtnext = dt_hyp;
          const double cap = C.cfl_max_var * dt;
          dtnext = dtnext <= cap ? dtnext : cap;
          // dt too small (main.c:534-543): flag it, take no more time
          if (dtnext < s.first_dt * 1.e-9) {
            s.dt_small = 1;
            dtnext = 0.0;
          }
          s.dt = dtnext;
        }
      }
    }
    __syncthreads();   // the patch is finished before its workgroup takes another
  }
}

// Stacked patches: before a step, each patch's scalars take its dt and
// start their reductions from zero.
__global__ void k_batch_pre(int nf, const int* ids, const PatchScalars* ps, char* heads, long hb)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nf) return;
  DevScalars* h = reinterpret_cast<DevScalars*>(heads + (long)q * hb);
  h->dt = ps[ids[q]].dt;
  h->invdt_max = 0ull;
  h->mach_max = 0ull;
  h->fail_cells = 0u;
  h->nan_cells = 0u;
  h->nan_next = 0u;
}

// ... and after it, the step is closed with k_step_end's rule for 3D boxes
// (NextTimeStep, main.c:459-543) into the patch's scalars.
__global__ void k_batch_post(int nf, const int* ids, PatchScalars* ps, const char* heads, long hb,
                             double cfl, double cfl_max_var, int fixed_dt)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nf) return;
  const DevScalars* h = reinterpret_cast<const DevScalars*>(heads + (long)q * hb);
  PatchScalars& s = ps[ids[q]];
  double inv = __longlong_as_double((long long)h->invdt_max);
  inv /= 3.0;
  const double mach = __longlong_as_double((long long)h->mach_max);
  const double dt = h->dt;
  s.invdt_last = inv;
  s.mach_last = mach;
  s.dt_last = dt;
  s.time += dt;
  s.step += 1;
  s.fail_cells += h->fail_cells;
  if (!fixed_dt) {
    double dt_hyp = 1.0 / inv;
    dt_hyp *= cfl;
    double dtnext = dt_hyp;
    const double cap = cfl_max_var * dt;
    dtnext = dtnext <= cap ? dtnext : cap;
    if (dtnext < s.first_dt * 1.e-9) {   // dt too small (main.c:534-543)
      s.dt_small = 1;
      dtnext = 0.0;
    }
    s.dt = dtnext;
  }
}

// flags of the stacked zones cleared (a kernel: no memset node in the graph)
__global__ void k_batch_clear_flags(unsigned char* flag, long n)
{
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x)
    flag[c] = 0;
}

// the step's faces, written to a third set by the fused stage 2 (its output
// may not alias its inputs), back into the state's faces
__global__ void k_batch_copy_faces(const double* c0, const double* c1, const double* c2, double* a0,
                                   double* a1, double* a2, long n)
{
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
    a0[c] = c0[c];
    a1[c] = c1[c];
    a2[c] = c2[c];
  }
}

}  // namespace pgpu

using namespace pgpu;

// an edge the stacked mode takes: the row kernel's and the LDS EMF's
// conditions (even, 4 to 256 cells)
static bool stack_edge(int n) { return n >= 4 && n <= 256 && n % 2 == 0; }

struct pgpu_batch {
  int npatch = 0;
  double gamma = 5.0 / 3.0;
  std::vector<int> edge;
  std::vector<long> off;
  long zones = 0;
  hipStream_t stream = nullptr;
  void* block = nullptr;
  size_t bytes = 0;
  BatchArrays A;
  BatchCtl C;
  PatchDesc* d_pd = nullptr;
  int* d_order = nullptr;
  PatchScalars* d_ps = nullptr;
  unsigned int* d_queue = nullptr;
  int threads = 256;     // workgroup size
  int groups = 0;        // resident workgroups
  int tile[3] = {0, 0, 0};   // cell tile of a pass (0: linear)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double cfl = 0.4, cfl_max_var = 1.1;
  int fixed_dt = 0;
  // stacked mode: the patches with stack_edge() edges come first in every
  // array, grouped by edge (stacks of at most 4 GiB per field)
  int mode = 1;
  struct Stack {
    int n, first, count;
    long off;
  };
  std::vector<Stack> stacks;
  int nfused = 0;                // patches in the stacks (positions 0 .. nfused-1)
  long zfused = 0;               // their zones, at the start of every array
  int nrest = 0;                 // patches of the persistent kernel in mode 1
  int* d_fused_ids = nullptr;    // [nfused] patch id of each stack position
  int* d_rest_order = nullptr;   // [nrest] the others, largest first
  char* d_heads = nullptr;       // [nfused] scalars (DevScalars before the histories)
  long head_bytes = 0;
  int stacked_now = 0;           // the captured step runs the stacks
};

// one stage of every stack through the fused kernels
static int enqueue_stacks(pgpu_batch* b, int stage)
{
  const BatchArrays& A = b->A;
  for (const auto& st : b->stacks) {
    const long o = st.off;
    Geo g;
    g.nx = g.ny = st.n;
    g.nz = st.count * st.n;
    g.dims = 3;
    g.n = (long)st.count * st.n * st.n * st.n;
    for (int a = 0; a < 3; a++) {
      g.dx[a] = 1.0 / st.n;
      g.inv_dx[a] = 1.0 / g.dx[a];
    }
    g.gamma = b->gamma;
    g.gmm1 = b->gamma - 1.0;
    g.pz = st.n;
    g.sc_stride = b->head_bytes;
    DevScalars* sc = reinterpret_cast<DevScalars*>(b->d_heads + (long)st.first * b->head_bytes);
    StageIn in, vn;
    for (int q = 0; q < 8; q++) {
      in.V[q] = (stage == 1 ? A.V[q] : A.V1[q]) + o;
      vn.V[q] = A.V[q] + o;
    }
    for (int a = 0; a < 3; a++) {
      in.b[a] = (stage == 1 ? A.bA[a] : A.bB[a]) + o;
      vn.b[a] = A.bA[a] + o;
    }
    Cons U;
    for (int q = 0; q < 5; q++) U.u[q] = A.U[q] + o;
    Sweeps sw;
    FaceData fd;
    for (int a = 0; a < 3; a++) {
      for (int q = 0; q < 5; q++) sw.d[a].F[q] = A.sw[a][q] + o;
      sw.d[a].P = A.sw[a][5] + o;
      sw.d[a].SL = fd.SL[a] = A.sw[a][6] + o;
      sw.d[a].SR = fd.SR[a] = A.sw[a][7] + o;
      sw.d[a].cm = A.sw[a][8] + o;
      for (int k = 0; k < 3; k++) sw.d[a].dv[k] = fd.dv[a][k] = A.sw[a][9 + k] + o;
    }
    double* cdt = A.sw[0][0] + o;   // (the flux arrays of the modular sweeps are free here)
    EdgeOut e;
    e.ex = A.e[0] + o;
    e.ey = A.e[1] + o;
    e.ez = A.e[2] + o;
    FaceIn fA, fB;
    FaceOut oB, oC;
    for (int a = 0; a < 3; a++) {
      fA.b[a] = A.bA[a] + o;
      fB.b[a] = A.bB[a] + o;
      oB.b[a] = A.bB[a] + o;
      oC.b[a] = A.sw[1][a] + o;   // third face set, copied back into bA after the step
    }
    Prims Vout;
    for (int q = 0; q < 8; q++) Vout.v[q] = (stage == 1 ? A.V1[q] : A.V[q]) + o;
    int rc = launch_flux_fused(g, stage, false, in, U, fd, cdt, sc, b->stream, nullptr);
    if (!rc) rc = launch_emf(g, in, sw, e, b->stream, nullptr);
    if (!rc)
      rc = stage == 1 ? launch_ct_c2p(g, 1, e, fA, fA, oB, U, vn, Vout, A.flag + o, sc, b->stream, nullptr)
                      : launch_ct_c2p(g, 2, e, fB, fA, oC, U, vn, Vout, A.flag + o, sc, b->stream, nullptr);
    if (rc) return rc;
  }
  return 0;
}

static int batch_capture(pgpu_batch* b)
{
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->C.cfl = b->cfl;
  b->C.cfl_max_var = b->cfl_max_var;
  b->C.fixed_dt = b->fixed_dt;
  b->C.tx = b->tile[0];
  b->C.ty = b->tile[1];
  b->C.tz = b->tile[2];
  // (the stacks need the row x/y kernel, which the process may have switched off)
  const bool stacked = b->mode == 1 && b->nfused > 0 && flux_xy_kind() == 1;
  b->stacked_now = stacked ? 1 : 0;
  // the persistent kernel's patches: all of them, or those outside the stacks
  BatchCtl C = b->C;
  if (stacked) {
    C.order = b->d_rest_order;
    C.npatch = b->nrest;
  }
  hipError_t err = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) return fail(PGPU_EDEVICE, "hipStreamBeginCapture: %s", hipGetErrorString(err));
  int rc = 0;
  auto persistent = [&](int stage) {
    if (C.npatch <= 0) return;
    if (b->threads <= 512) {
      if (stage == 1) hipLaunchKernelGGL((k_batch_stage<1, 512>), dim3(b->groups), dim3(b->threads), 0, b->stream, b->A, C);
      else hipLaunchKernelGGL((k_batch_stage<2, 512>), dim3(b->groups), dim3(b->threads), 0, b->stream, b->A, C);
    } else {
      if (stage == 1) hipLaunchKernelGGL((k_batch_stage<1, 1024>), dim3(b->groups), dim3(b->threads), 0, b->stream, b->A, C);
      else hipLaunchKernelGGL((k_batch_stage<2, 1024>), dim3(b->groups), dim3(b->threads), 0, b->stream, b->A, C);
    }
  };
  const int nfb = (b->nfused + 255) / 256;
  if (stacked) {
    hipLaunchKernelGGL(k_batch_pre, dim3(nfb), dim3(256), 0, b->stream, b->nfused, b->d_fused_ids,
                       b->d_ps, b->d_heads, b->head_bytes);
    hipLaunchKernelGGL(k_batch_clear_flags, dim3(2048), dim3(256), 0, b->stream, b->A.flag, b->zfused);
  }
  persistent(1);
  if (stacked) rc = enqueue_stacks(b, 1);
  persistent(2);
  if (stacked && !rc) rc = enqueue_stacks(b, 2);
  if (stacked && !rc) {
    hipLaunchKernelGGL(k_batch_copy_faces, dim3(4096), dim3(256), 0, b->stream, b->A.sw[1][0], b->A.sw[1][1],
                       b->A.sw[1][2], b->A.bA[0], b->A.bA[1], b->A.bA[2], b->zfused);
    hipLaunchKernelGGL(k_batch_post, dim3(nfb), dim3(256), 0, b->stream, b->nfused, b->d_fused_ids,
                       b->d_ps, b->d_heads, b->head_bytes, b->cfl, b->cfl_max_var, b->fixed_dt);
  }
  err = hipStreamEndCapture(b->stream, &b->graph);
  if (rc) return fail(PGPU_EDEVICE, "stacked stage launch failed (%d)", rc);
  if (err != hipSuccess) return fail(PGPU_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(err));
  err = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (err != hipSuccess) return fail(PGPU_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(err));
  return PGPU_OK;
}

static int default_groups(const pgpu_batch* b)
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
  }
  const int g = cus * 4;
  return g < b->npatch ? g : b->npatch;
}

extern "C" int pgpu_batch_create(int npatch, const int* edge, double gamma, pgpu_batch** out)
{
  if (!out || !edge || npatch <= 0) return fail(PGPU_EINVAL, "bad batch arguments");
  *out = nullptr;
  pgpu_batch* b = new pgpu_batch;
  b->npatch = npatch;
  b->gamma = gamma;
  b->edge.assign(edge, edge + npatch);
  b->off.resize(npatch);
  for (int p = 0; p < npatch; p++) {
    const int n = edge[p];
    if (n < 3 || n > 1024) {
      delete b;
      return fail(PGPU_EINVAL, "patch %d: edge %d outside [3, 1024]", p, n);
    }
  }
  // storage order: the stackable patches by edge (ascending), in index order
  // within an edge, then the others in index order
  std::vector<int> pos(npatch);
  std::iota(pos.begin(), pos.end(), 0);
  std::stable_sort(pos.begin(), pos.end(), [&](int x, int y) {
    const bool sx = stack_edge(edge[x]), sy = stack_edge(edge[y]);
    if (sx != sy) return sx;
    return sx && edge[x] < edge[y];
  });
  long z = 0;
  std::vector<int> fused_ids, rest;
  for (int q = 0; q < npatch; q++) {
    const int p = pos[q], n = edge[p];
    const long cells = (long)n * n * n;
    b->off[p] = z;
    if (stack_edge(n)) {
      // a new stack at a new edge, or where the stack would reach 4 GiB per field
      if (b->stacks.empty() || b->stacks.back().n != n ||
          (long)(b->stacks.back().count + 1) * cells * 8 >= (1l << 32))
        b->stacks.push_back({n, (int)fused_ids.size(), 0, z});
      b->stacks.back().count++;
      fused_ids.push_back(p);
      b->zfused = z + cells;
    } else {
      rest.push_back(p);
    }
    z += cells;
  }
  b->nfused = (int)fused_ids.size();
  b->nrest = (int)rest.size();
  std::stable_sort(rest.begin(), rest.end(), [&](int x, int y) { return edge[x] > edge[y]; });
  b->zones = z;
  // 8 + 8 + 5 + 3 + 3 state arrays, 36 sweep arrays, 3 edges
  const int narr = 8 + 8 + 5 + 3 + 3 + 36 + 3;
  const size_t per = (size_t)z * sizeof(double);
  const size_t flag_b = ((size_t)z + 255) & ~(size_t)255;
  b->bytes = per * narr + flag_b;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(PGPU_EDEVICE, "hipStreamCreate failed");
  }
  if (hipMalloc(&b->block, b->bytes) != hipSuccess) {
    const size_t bytes = b->bytes;
    (void)hipStreamDestroy(b->stream);
    delete b;
    return fail(PGPU_ENOMEM, "hipMalloc of %zu bytes for the batch failed", bytes);
  }
  char* cur = (char*)b->block;
  auto take = [&]() {
    double* p = (double*)cur;
    cur += per;
    return p;
  };
  BatchArrays& A = b->A;
  for (int q = 0; q < 8; q++) A.V[q] = take();
  for (int q = 0; q < 8; q++) A.V1[q] = take();
  for (int q = 0; q < 5; q++) A.U[q] = take();
  for (int a = 0; a < 3; a++) A.bA[a] = take();
  for (int a = 0; a < 3; a++) A.bB[a] = take();
  for (int dd = 0; dd < 3; dd++)
    for (int k = 0; k < 12; k++) A.sw[dd][k] = take();
  for (int a = 0; a < 3; a++) A.e[a] = take();
  A.flag = (unsigned char*)cur;

  std::vector<PatchDesc> pd(npatch);
  std::vector<PatchScalars> ps(npatch);
  for (int p = 0; p < npatch; p++) {
    pd[p].n = edge[p];
    pd[p].pad = 0;
    pd[p].off = b->off[p];
    pd[p].dx = 1.0 / edge[p];
    ps[p] = PatchScalars{};
    ps[p].dt = 1e-4;
  }
  std::vector<int> order(npatch);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return edge[x] > edge[y]; });
  hipError_t err = hipMalloc(&b->d_pd, sizeof(PatchDesc) * npatch);
  if (err == hipSuccess) err = hipMalloc(&b->d_order, sizeof(int) * npatch);
  if (err == hipSuccess) err = hipMalloc(&b->d_ps, sizeof(PatchScalars) * npatch);
  if (err == hipSuccess) err = hipMalloc(&b->d_queue, 2 * sizeof(unsigned int));
  if (err == hipSuccess) err = hipMemcpy(b->d_pd, pd.data(), sizeof(PatchDesc) * npatch, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(b->d_order, order.data(), sizeof(int) * npatch, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(b->d_ps, ps.data(), sizeof(PatchScalars) * npatch, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemset(b->d_queue, 0, 2 * sizeof(unsigned int));
  if (err == hipSuccess) err = hipMemset(A.flag, 0, flag_b);
  b->head_bytes = ((long)offsetof(DevScalars, dt_hist) + 127) / 128 * 128;
  if (err == hipSuccess && b->nfused) {
    err = hipMalloc(&b->d_fused_ids, sizeof(int) * b->nfused);
    if (err == hipSuccess) err = hipMalloc(&b->d_heads, (size_t)b->head_bytes * b->nfused);
    if (err == hipSuccess)
      err = hipMemcpy(b->d_fused_ids, fused_ids.data(), sizeof(int) * b->nfused, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(b->d_heads, 0, (size_t)b->head_bytes * b->nfused);
  }
  if (err == hipSuccess && b->nrest) {
    err = hipMalloc(&b->d_rest_order, sizeof(int) * b->nrest);
    if (err == hipSuccess)
      err = hipMemcpy(b->d_rest_order, rest.data(), sizeof(int) * b->nrest, hipMemcpyHostToDevice);
  }
  if (err != hipSuccess) {
    pgpu_batch_destroy(b);
    return fail(PGPU_EDEVICE, "batch set-up failed: %s", hipGetErrorString(err));
  }
  b->C.pd = b->d_pd;
  b->C.order = b->d_order;
  b->C.ps = b->d_ps;
  b->C.queue = b->d_queue;
  b->C.npatch = npatch;
  b->C.gamma = gamma;
  b->C.gmm1 = gamma - 1.0;
  b->groups = default_groups(b);
  int rc = batch_capture(b);
  if (rc) {
    pgpu_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PGPU_OK;
}

extern "C" int pgpu_batch_destroy(pgpu_batch* b)
{
  if (!b) return PGPU_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->block) (void)hipFree(b->block);
  if (b->d_pd) (void)hipFree(b->d_pd);
  if (b->d_order) (void)hipFree(b->d_order);
  if (b->d_ps) (void)hipFree(b->d_ps);
  if (b->d_queue) (void)hipFree(b->d_queue);
  if (b->d_fused_ids) (void)hipFree(b->d_fused_ids);
  if (b->d_rest_order) (void)hipFree(b->d_rest_order);
  if (b->d_heads) (void)hipFree(b->d_heads);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PGPU_OK;
}

static bool patch_ok(const pgpu_batch* b, int p) { return b && p >= 0 && p < b->npatch; }

extern "C" int pgpu_batch_set_patch(pgpu_batch* b, int p, const double* prim, const double* bx,
                                    const double* by, const double* bz)
{
  if (!patch_ok(b, p) || !prim || !bx || !by || !bz) return fail(PGPU_EINVAL, "bad patch arguments");
  if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(PGPU_EDEVICE, "stream synchronisation failed");
  const int n = b->edge[p];
  const size_t nz = (size_t)n * n * n, nb = nz * sizeof(double);
  const long o = b->off[p];
  // the cell-centred field is the face average (as pgpu_set_state)
  std::vector<double> cb(nz);
  const double* f[3] = {bx, by, bz};
  hipError_t err = hipSuccess;
  for (int q = 0; q < 8 && err == hipSuccess; q++) {
    if (q >= BX1 && q <= BX3) {
      const int a = q - BX1;
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) {
            const size_t c = ((size_t)k * n + j) * n + i;
            const int im = a == 0 ? (i + n - 1) % n : i;
            const int jm = a == 1 ? (j + n - 1) % n : j;
            const int km = a == 2 ? (k + n - 1) % n : k;
            cb[c] = 0.5 * (f[a][c] + f[a][((size_t)km * n + jm) * n + im]);
          }
      err = hipMemcpy(b->A.V[q] + o, cb.data(), nb, hipMemcpyHostToDevice);
    } else {
      err = hipMemcpy(b->A.V[q] + o, prim + (size_t)q * nz, nb, hipMemcpyHostToDevice);
    }
  }
  for (int a = 0; a < 3 && err == hipSuccess; a++) err = hipMemcpy(b->A.bA[a] + o, f[a], nb, hipMemcpyHostToDevice);
  if (err != hipSuccess) return fail(PGPU_EDEVICE, "patch upload failed: %s", hipGetErrorString(err));
  return PGPU_OK;
}

extern "C" int pgpu_batch_get_patch(pgpu_batch* b, int p, double* prim, double* bx, double* by,
                                    double* bz)
{
  if (!patch_ok(b, p)) return fail(PGPU_EINVAL, "bad patch index");
  if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(PGPU_EDEVICE, "stream synchronisation failed");
  const int n = b->edge[p];
  const size_t nz = (size_t)n * n * n, nb = nz * sizeof(double);
  const long o = b->off[p];
  hipError_t err = hipSuccess;
  if (prim)
    for (int q = 0; q < 8 && err == hipSuccess; q++)
      err = hipMemcpy(prim + (size_t)q * nz, b->A.V[q] + o, nb, hipMemcpyDeviceToHost);
  double* f[3] = {bx, by, bz};
  for (int a = 0; a < 3 && err == hipSuccess; a++)
    if (f[a]) err = hipMemcpy(f[a], b->A.bA[a] + o, nb, hipMemcpyDeviceToHost);
  if (err != hipSuccess) return fail(PGPU_EDEVICE, "patch download failed: %s", hipGetErrorString(err));
  return PGPU_OK;
}

extern "C" int pgpu_batch_set_dt(pgpu_batch* b, double dt, double cfl, double cfl_max_var, int fixed_dt)
{
  if (!b) return fail(PGPU_EINVAL, "batch is NULL");
  if (!(dt > 0.0)) return fail(PGPU_EINVAL, "dt must be positive");
  if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(PGPU_EDEVICE, "stream synchronisation failed");
  std::vector<PatchScalars> ps(b->npatch);
  for (auto& s : ps) {
    s = PatchScalars{};
    s.dt = dt;
    s.first_dt = dt;
  }
  if (hipMemcpy(b->d_ps, ps.data(), sizeof(PatchScalars) * b->npatch, hipMemcpyHostToDevice) != hipSuccess)
    return fail(PGPU_EDEVICE, "dt upload failed");
  b->cfl = cfl;
  b->cfl_max_var = cfl_max_var;
  b->fixed_dt = fixed_dt ? 1 : 0;
  return batch_capture(b);
}

extern "C" int pgpu_batch_run(pgpu_batch* b, int nsteps)
{
  if (!b) return fail(PGPU_EINVAL, "batch is NULL");
  for (int s = 0; s < nsteps; s++) {
    const hipError_t err = hipGraphLaunch(b->exec, b->stream);
    if (err != hipSuccess) return fail(PGPU_EDEVICE, "hipGraphLaunch: %s", hipGetErrorString(err));
  }
  return PGPU_OK;
}

extern "C" int pgpu_batch_sync(pgpu_batch* b)
{
  if (!b) return fail(PGPU_EINVAL, "batch is NULL");
  if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(PGPU_EDEVICE, "stream synchronisation failed");
  std::vector<PatchScalars> ps(b->npatch);
  if (hipMemcpy(ps.data(), b->d_ps, sizeof(PatchScalars) * b->npatch, hipMemcpyDeviceToHost) !=
      hipSuccess)
    return fail(PGPU_EDEVICE, "scalar download failed");
  for (int p = 0; p < b->npatch; p++)
    if (ps[p].dt_small)
      return fail(PGPU_EDTSMALL, "patch %d: dt is too small (< 1e-9 of its first dt); PLUTO's "
                                 "NextTimeStep stops the run there", p);
  return PGPU_OK;
}

extern "C" int pgpu_batch_time_steps(pgpu_batch* b, int nsteps, double* ms)
{
  if (!b || !ms) return fail(PGPU_EINVAL, "NULL argument");
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(PGPU_EDEVICE, "hipEventCreate failed");
  (void)hipEventRecord(e0, b->stream);
  int rc = pgpu_batch_run(b, nsteps);
  (void)hipEventRecord(e1, b->stream);
  float t = 0.f;
  if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess))
    rc = fail(PGPU_EDEVICE, "event timing failed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t;
  return rc;
}

extern "C" int pgpu_batch_patch_scalars(pgpu_batch* b, int p, double out[6])
{
  if (!patch_ok(b, p) || !out) return fail(PGPU_EINVAL, "bad patch arguments");
  PatchScalars s;
  if (hipStreamSynchronize(b->stream) != hipSuccess ||
      hipMemcpy(&s, b->d_ps + p, sizeof(s), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(PGPU_EDEVICE, "patch scalar download failed");
  out[0] = s.dt;
  out[1] = s.invdt_last;
  out[2] = s.mach_last;
  out[3] = s.time;
  out[4] = (double)s.step;
  out[5] = (double)s.fail_cells;
  return PGPU_OK;
}

extern "C" int pgpu_batch_set_launch(pgpu_batch* b, int threads, int groups)
{
  if (!b) return fail(PGPU_EINVAL, "batch is NULL");
  if (threads != 64 && threads != 128 && threads != 256 && threads != 512 && threads != 1024)
    return fail(PGPU_EINVAL, "workgroup size %d not in {64, 128, 256, 512, 1024}", threads);
  if (groups <= 0) groups = default_groups(b);
  if (groups > b->npatch) groups = b->npatch;
  b->threads = threads;
  b->groups = groups;
  if (b->tile[0] && b->tile[0] * b->tile[1] * b->tile[2] != threads)
    b->tile[0] = b->tile[1] = b->tile[2] = 0;   // the tile no longer fits: linear
  return batch_capture(b);
}

extern "C" int pgpu_batch_set_tile(pgpu_batch* b, int tx, int ty, int tz)
{
  if (!b) return fail(PGPU_EINVAL, "batch is NULL");
  if (tx == 0 && ty == 0 && tz == 0) {
    b->tile[0] = b->tile[1] = b->tile[2] = 0;
    return batch_capture(b);
  }
  if (tx < 1 || ty < 1 || tz < 1 || tx * ty * tz != b->threads)
    return fail(PGPU_EINVAL, "tile %d x %d x %d must hold exactly the workgroup's %d threads",
                tx, ty, tz, b->threads);
  b->tile[0] = tx;
  b->tile[1] = ty;
  b->tile[2] = tz;
  return batch_capture(b);
}

extern "C" int pgpu_batch_set_mode(pgpu_batch* b, int mode)
{
  if (!b) return fail(PGPU_EINVAL, "batch is NULL");
  if (mode != 0 && mode != 1) return fail(PGPU_EINVAL, "batch mode %d: 0 (persistent) or 1 (stacked)", mode);
  if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(PGPU_EDEVICE, "stream synchronisation failed");
  b->mode = mode;
  return batch_capture(b);
}

extern "C" int pgpu_batch_get_mode(pgpu_batch* b, int* mode, int* stacked_patches)
{
  if (!b || !mode) return fail(PGPU_EINVAL, "NULL argument");
  *mode = b->mode;
  if (stacked_patches) *stacked_patches = b->stacked_now ? b->nfused : 0;
  return PGPU_OK;
}

extern "C" int pgpu_batch_get_tile(pgpu_batch* b, int* tx, int* ty, int* tz)
{
  if (!b || !tx || !ty || !tz) return fail(PGPU_EINVAL, "NULL argument");
  *tx = b->tile[0];
  *ty = b->tile[1];
  *tz = b->tile[2];
  return PGPU_OK;
}

extern "C" int pgpu_batch_get_launch(pgpu_batch* b, int* threads, int* groups)
{
  if (!b || !threads || !groups) return fail(PGPU_EINVAL, "NULL argument");
  *threads = b->threads;
  *groups = b->groups;
  return PGPU_OK;
}

extern "C" int pgpu_batch_bytes(pgpu_batch* b, size_t* bytes)
{
  if (!b || !bytes) return fail(PGPU_EINVAL, "NULL argument");
  *bytes = b->bytes;
  return PGPU_OK;
}

// Autotuner: times every launch shape (workgroup size x resident
// workgroups per CU) with the linear cell order, then every cell tile of
// the fastest workgroup size (boxes of its threads, 4 to 64 cells along x),
// each over trial_steps graph replays, and keeps the fastest.  The trial
// steps advance the patches like ordinary steps.
extern "C" int pgpu_batch_tune(pgpu_batch* b, int trial_steps, int* threads, int* groups,
                               double* ms_per_step)
{
  if (!b || trial_steps <= 0) return fail(PGPU_EINVAL, "bad tuning arguments");
  if (b->stacked_now && b->nrest == 0) {
    // every patch is stacked: the launch shape does not apply; time the step
    double ms = 0.0;
    int rc = pgpu_batch_run(b, 1);
    if (!rc) rc = pgpu_batch_time_steps(b, trial_steps, &ms);
    if (rc) return rc;
    if (threads) *threads = 0;
    if (groups) *groups = 0;
    if (ms_per_step) *ms_per_step = ms / trial_steps;
    return PGPU_OK;
  }
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
  }
  const int tsizes[4] = {128, 256, 512, 1024};
  const int per_cu[4] = {1, 2, 4, 8};
  double best = 1e300;
  int bt = b->threads, bg = b->groups;
  b->tile[0] = b->tile[1] = b->tile[2] = 0;
  for (int t : tsizes) {
    for (int k : per_cu) {
      if (t * k > 2048) continue;   // beyond one CU's resident threads
      int rc = pgpu_batch_set_launch(b, t, cus * k);
      if (!rc) rc = pgpu_batch_run(b, 1);
      double ms = 0.0;
      if (!rc) rc = pgpu_batch_time_steps(b, trial_steps, &ms);
      if (rc) return rc;
      ms /= trial_steps;
      if (ms < best) {
        best = ms;
        bt = b->threads;
        bg = b->groups;
      }
    }
  }
  int rc = pgpu_batch_set_launch(b, bt, bg);
  if (rc) return rc;
  // cell tiles of bt threads: tx in {4..64} (powers of two), ty >= tz
  int btile[3] = {0, 0, 0};
  for (int tx = 4; tx <= 64; tx *= 2) {
    if (bt % tx) continue;
    const int rest = bt / tx;
    for (int tz = 1; tz * tz <= rest; tz *= 2) {
      if (rest % tz) continue;
      const int ty = rest / tz;
      rc = pgpu_batch_set_tile(b, tx, ty, tz);
      if (!rc) rc = pgpu_batch_run(b, 1);
      double ms = 0.0;
      if (!rc) rc = pgpu_batch_time_steps(b, trial_steps, &ms);
      if (rc) return rc;
      ms /= trial_steps;
      if (ms < best) {
        best = ms;
        btile[0] = tx;
        btile[1] = ty;
        btile[2] = tz;
      }
    }
  }
  rc = pgpu_batch_set_tile(b, btile[0], btile[1], btile[2]);
  if (rc) return rc;
  if (threads) *threads = bt;
  if (groups) *groups = bg;
  if (ms_per_step) *ms_per_step = best;
  return PGPU_OK;
}
