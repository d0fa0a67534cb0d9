// Internal definition of a domain, shared by the native C-ABI (api.hip)
// and PLUTO's entry points (pluto_abi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/pluto_amd.h"
#include "boundary_kernels.hpp"
#include "stage_kernels.hpp"

namespace pgpu {

// Records a formatted error for pgpu_last_error() and returns code.
int fail(int code, const char* fmt, ...);

// ---------------------------------------------------------------------
// kernel timer
// ---------------------------------------------------------------------


struct EventTimer final : KernelTimer {
  struct Rec {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t pool_next = 0;
  double ms[KT_COUNT] = {0};
  long count[KT_COUNT] = {0};
  hipEvent_t pending_begin = nullptr;

  hipEvent_t take()
  {
    if (pool_next == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[pool_next++];
  }
  void begin(int id, hipStream_t s) override
  {
    (void)id;
    pending_begin = take();
    if (pending_begin) (void)hipEventRecord(pending_begin, s);
  }
  void end(int id, hipStream_t s) override
  {
    hipEvent_t e = take();
    if (!e || !pending_begin) return;
    (void)hipEventRecord(e, s);
    recs.push_back({id, pending_begin, e});
    pending_begin = nullptr;
  }
  // Fold recorded pairs into totals (caller synchronised the stream).
  void collect()
  {
    for (const Rec& r : recs) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
        ms[r.id] += t;
        count[r.id] += 1;
      }
    }
    recs.clear();
    pool_next = 0;
  }
  ~EventTimer() override
  {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};


}  // namespace pgpu

struct pgpu_adjoint_state;   // adjoint work arrays (adjoint_kernels.hip)

struct pgpu_domain {
  pgpu::Geo g;
  int device = 0;
  int precision = PGPU_PREC_FP64;
  hipStream_t stream = nullptr;
  void* block = nullptr;   // one allocation for every array
  size_t bytes = 0;
  double* V[8];            // step-start primitives, overwritten by the step result
  double* V1[8];           // stage-1 primitives
  double* U[5];            // conservative accumulator
  double* bA[3];           // faces: step start, then step result
  double* bB[3];           // faces after stage 1
  double* bC[3];           // stage-2 face output of the fused pipeline (rotated into bA)
  double* cdt;             // C_dt, x+y part (fused pipeline)
  int pipeline = PGPU_PIPELINE_FUSED;
  pgpu::Sweeps sw;
  pgpu::EdgeOut e;
  unsigned char* flag = nullptr;
  pgpu::DevScalars* sc = nullptr;
  double cfl = 0.4, cfl_max_var = 1.1;
  int fixed_dt = 0;
  pgpu::EventTimer* timer = nullptr;
  pgpu_adjoint_state* adj = nullptr;
  // ghosted box (g.ghost): the boundary types of the six sides, the layout
  // and the face areas the ghost fill between stages uses
  int bound[6] = {0, 0, 0, 0, 0, 0};
  pgpu::GhostBox gb;
  double area[3] = {0.0, 0.0, 0.0};
  // branch signature of the steps run while enabled (null: off)
  unsigned long long* sig = nullptr;
  bool sig_on = false;
  // branch tape (mode 0 off, 1 record, 2 replay): 4 words per cell and
  // stage for tape_cap steps; tape_step the next step, tape_len recorded
  unsigned long long* tape = nullptr;
  int tape_mode = 0, tape_cap = 0, tape_step = 0, tape_len = 0;
  // USERDEF sides: the host hook and its slab staging
  pgpu_userdef_fn userdef = nullptr;
  void* userdef_ctx = nullptr;
  std::vector<double> slab_buf;
};

// Makes a domain of (t[0]+1) x (t[1]+1) x (t[2]+1) cells hold PLUTO's
// ghosted box (t cells per axis, domain beg..end, at offset +1: see
// pgpu::GhostBox) with boundary types type[6] (OUTFLOW .. PERIODIC for
// X1_BEG, X1_END, ... X3_END) and uniform face areas area[3]: every stage
// starts with Boundary()'s ghost fill, and the CFL, Mach, failure and NaN
// reductions keep to PLUTO's ranges.  Fused FP64 pipeline only.
int pgpu_domain_set_ghosted(pgpu_domain* d, const int t[3], const int beg[3], const int end[3],
                            const int type[6], const double area[3]);

// A bounded box's state from interior arrays, with the face planes below the
// first cells given (low: bx [nz][ny], by [nz][nx], bz [ny][nx]) or, when
// low is null, the images of the last faces.
int pgpu_ghosted_set_state(pgpu_domain* d, const double* prim, const double* const faces[3],
                           const double* const* low);
// The face planes below the first cells of a bounded box (layout of low above).
int pgpu_ghosted_low_faces(pgpu_domain* d, double* const* low);

// Releases the adjoint work arrays of a domain (no-op when none).
void pgpu_adjoint_free(pgpu_domain* d);
