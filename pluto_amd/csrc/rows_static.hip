// pgpu_riemann_rows_static: the Riemann-solver row entry point (the body of
// HLLD_Solver, HLL_Solver, HLLC_Solver, LF_Solver and Roe_Solver) for
// callers that cannot allocate device memory.  The row entry points of
// pluto_abi.hip grow a device buffer with hipMalloc on demand; this variant
// stages the rows through `scratch`, a fixed device array of this file, and
// fixed host arrays, and calls no allocating routine: no hipMalloc/hipFree,
// no stream creation (it runs on the null stream), no heap allocation.
//
// Every call on a GPU shares that GPU's copy of `scratch` (and the process's
// host staging arrays): calls that overlap, from several host threads or
// with other work that uses the same rows, race.  Each call writes every
// scratch element it reads, so nothing must start at zero.

#include <hip/hip_runtime.h>
#include <stdio.h>

#include "../../include/pluto_abi.h"
#include "../../include/pluto_amd.h"
#include "mhd_device.hpp"
#include "row_kernels.hpp"

using namespace pgpu;

extern "C" int g_dir;
extern "C" double g_gamma;
extern "C" double g_maxMach;

namespace {

constexpr int NVS = 8;                    // NVAR of the 3D MHD build
constexpr int PER_ROW = 3 * NVS + 5;      // vL, vR, flux; press, SL, SR, cmax, Mach
constexpr size_t SCRATCH_DOUBLES = (size_t)PGPU_STATIC_ROWS_MAX * PER_ROW;
constexpr int MAX_DEVICES = 64;

// host staging, same layout as the device scratch
double h_stage[SCRATCH_DOUBLES];

// per-device address of `scratch`, from hipGetSymbolAddress
void* dev_ptr[MAX_DEVICES];

}  // namespace

// 8192 rows x 29 doubles = 1.9 MB per GPU, held until the process ends
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {

int scratch_ptr(double** out)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    fprintf(stderr, "! pgpu_riemann_rows_static (pluto_amd): no usable HIP device\n");
    return PGPU_EDEVICE;
  }
  if (!dev_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "! pgpu_riemann_rows_static (pluto_amd): hipGetSymbolAddress: %s\n",
              hipGetErrorString(e));
      return PGPU_EDEVICE;
    }
    dev_ptr[dev] = p;
  }
  *out = (double*)dev_ptr[dev];
  return PGPU_OK;
}

}  // namespace

extern "C" int pgpu_riemann_rows_static(int solver, const Sweep* sweep, int beg, int end,
                                        double* cmax)
{
  if (!sweep || !cmax) return PGPU_EINVAL;
  if (solver < PGPU_SOLVER_HLLD || solver > PGPU_SOLVER_ROE) {
    fprintf(stderr, "! pgpu_riemann_rows_static (pluto_amd): unknown solver %d\n", solver);
    return PGPU_EINVAL;
  }
  const int n = end - beg + 1;
  if (n <= 0) return PGPU_OK;
  if (n > PGPU_STATIC_ROWS_MAX) {
    fprintf(stderr,
            "! pgpu_riemann_rows_static (pluto_amd): %d interfaces exceed the fixed scratch "
            "of %d rows; nothing was computed\n",
            n, PGPU_STATIC_ROWS_MAX);
    return PGPU_EINVAL;
  }
  double* d = nullptr;
  const int rc = scratch_ptr(&d);
  if (rc != PGPU_OK) return rc;

  // same packing as the allocating entry point: vL rows, vR rows, then the
  // outputs flux rows, press, SL, SR, cmax, Mach
  double *dvl = d, *dvr = d + (size_t)n * NVS, *dfl = d + 2 * (size_t)n * NVS;
  double *dp = d + 3 * (size_t)n * NVS, *dsl = dp + n, *dsr = dsl + n, *dcm = dsr + n,
         *dmach = dcm + n;
  double* h = h_stage;
  for (int i = 0; i < n; i++)
    for (int q = 0; q < NVS; q++) {
      h[(size_t)i * NVS + q] = sweep->stateL.v[beg + i][q];
      h[(size_t)(n + i) * NVS + q] = sweep->stateR.v[beg + i][q];
    }
  const hipStream_t s = 0;
  bool ok = hipMemcpyAsync(d, h, 2 * (size_t)n * NVS * sizeof(double), hipMemcpyHostToDevice,
                           s) == hipSuccess;
  ok = ok && launch_riemann_rows(solver, g_dir, n, dvl, dvr, g_gamma, dfl, dp, dsl, dsr, dcm,
                                 dmach, s) == 0;
  ok = ok && hipMemcpyAsync(h + 2 * (size_t)n * NVS, dfl,
                            ((size_t)n * NVS + 5 * (size_t)n) * sizeof(double),
                            hipMemcpyDeviceToHost, s) == hipSuccess;
  ok = ok && hipStreamSynchronize(s) == hipSuccess;
  if (!ok) {
    fprintf(stderr, "! pgpu_riemann_rows_static (pluto_amd): device launch failed\n");
    return PGPU_EDEVICE;
  }
  const double* f = h + 2 * (size_t)n * NVS;
  const double* r = f + (size_t)n * NVS;
  for (int i = 0; i < n; i++) {
    for (int q = 0; q < NVS; q++) sweep->flux[beg + i][q] = f[(size_t)i * NVS + q];
    sweep->press[beg + i] = r[i];
    sweep->SL[beg + i] = r[n + i];
    sweep->SR[beg + i] = r[2 * n + i];
    cmax[beg + i] = r[3 * n + i];
    const double m = r[4 * n + i];
    g_maxMach = (m >= g_maxMach) ? m : g_maxMach;   // MAX(scrh, g_maxMach)
  }
  return PGPU_OK;
}
