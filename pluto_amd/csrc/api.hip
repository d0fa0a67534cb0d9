// Host library: domains resident in device memory and the native C-ABI
// declared in include/pluto_amd.h.

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "../../include/pluto_amd.h"
#include "mhd_device.hpp"
#include "stage_kernels.hpp"
#include "domain.hpp"
#include "fused_kernels.hpp"

using namespace pgpu;

// ---------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------

static thread_local char g_err[512] = "";

int pgpu::fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(call)                                                              \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess)                                                          \
      return fail(PGPU_EDEVICE, "%s failed: %s", #call, hipGetErrorString(_e));   \
  } while (0)

extern "C" const char* pgpu_last_error(void) { return g_err; }

extern "C" int pgpu_device_count(int* count)
{
  if (!count) return fail(PGPU_EINVAL, "count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return PGPU_OK;
}

static const char* kKernelNames[KT_COUNT] = {"sweep_x", "sweep_y", "sweep_z",  "update",
                                             "emf",     "ct",      "c2p",      "step_end",
                                             "flux_xy", "flux_z",  "boundary"};

extern "C" int pgpu_kernel_name(int id, const char** name)
{
  if (id < 0 || id >= KT_COUNT || !name) return fail(PGPU_EINVAL, "bad kernel id %d", id);
  *name = kKernelNames[id];
  return PGPU_OK;
}

extern "C" int pgpu_domain_create(int nx, int ny, int nz, int dims, double dx, double dy,
                                  double dz, double gamma, pgpu_domain** out)
{
  if (!out) return fail(PGPU_EINVAL, "out is NULL");
  *out = nullptr;
  if (dims != 2 && dims != 3) return fail(PGPU_EINVAL, "dims must be 2 or 3 (got %d)", dims);
  if (nx < 3 || ny < 3) return fail(PGPU_EINVAL, "nx, ny must be >= 3 (got %d, %d)", nx, ny);
  if (dims == 3 && nz < 3) return fail(PGPU_EINVAL, "3D boxes need nz >= 3 (got %d)", nz);
  if (dims == 2 && nz != 1) return fail(PGPU_EINVAL, "2D boxes need nz == 1 (got %d)", nz);
  if (!(dx > 0.0) || !(dy > 0.0) || !(dz > 0.0))
    return fail(PGPU_EINVAL, "cell widths must be positive");
  if (!(gamma > 1.0)) return fail(PGPU_EINVAL, "gamma must exceed 1");
  if (nz > 65535) return fail(PGPU_EINVAL, "nz exceeds the launch grid limit");

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(PGPU_EDEVICE, "no HIP device available");

  pgpu_domain* d = new (std::nothrow) pgpu_domain();
  if (!d) return fail(PGPU_ENOMEM, "host allocation failed");
  Geo& g = d->g;
  g.nx = nx; g.ny = ny; g.nz = nz; g.dims = dims;
  g.n = (long)nx * ny * nz;
  g.dx[0] = dx; g.dx[1] = dy; g.dx[2] = dz;
  for (int a = 0; a < 3; a++) g.inv_dx[a] = 1.0 / g.dx[a];
  g.gamma = gamma;
  g.gmm1 = gamma - 1.0;

  (void)hipGetDevice(&d->device);
  // arrays: V 8, V1 8, U 5, bA 3, bB 3, bC 3, per direction 12, edges 3, C_dt 1
  const int narr = 8 + 8 + 5 + 3 + 3 + 3 + 3 * 12 + 3 + 1;
  const size_t per = ((size_t)g.n * sizeof(double) + 255) & ~(size_t)255;
  const size_t flag_bytes = ((size_t)g.n + 255) & ~(size_t)255;
  const size_t sc_bytes = (sizeof(DevScalars) + 255) & ~(size_t)255;
  d->bytes = per * narr + flag_bytes + sc_bytes;
  hipError_t err = hipMalloc(&d->block, d->bytes);
  if (err != hipSuccess) {
    const size_t bytes = d->bytes;
    delete d;
    return fail(PGPU_ENOMEM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(err));
  }
  char* p = (char*)d->block;
  auto take = [&](void) -> double* {
    double* r = (double*)p;
    p += per;
    return r;
  };
  for (int q = 0; q < 8; q++) d->V[q] = take();
  for (int q = 0; q < 8; q++) d->V1[q] = take();
  for (int q = 0; q < 5; q++) d->U[q] = take();
  for (int q = 0; q < 3; q++) d->bA[q] = take();
  for (int q = 0; q < 3; q++) d->bB[q] = take();
  for (int q = 0; q < 3; q++) d->bC[q] = take();
  d->cdt = take();
  for (int a = 0; a < 3; a++) {
    SweepOut& s = d->sw.d[a];
    for (int q = 0; q < 5; q++) s.F[q] = take();
    s.P = take();
    s.SL = take();
    s.SR = take();
    s.cm = take();
    for (int q = 0; q < 3; q++) s.dv[q] = take();
  }
  d->e.ex = take();
  d->e.ey = take();
  d->e.ez = take();
  d->flag = (unsigned char*)p;
  p += flag_bytes;
  d->sc = (DevScalars*)p;

  if (hipStreamCreate(&d->stream) != hipSuccess) {
    (void)hipFree(d->block);
    delete d;
    return fail(PGPU_EDEVICE, "hipStreamCreate failed");
  }
  // zero everything once so unused planes (2D) hold defined values
  if (hipMemsetAsync(d->block, 0, d->bytes, d->stream) != hipSuccess ||
      hipStreamSynchronize(d->stream) != hipSuccess) {
    (void)hipStreamDestroy(d->stream);
    (void)hipFree(d->block);
    delete d;
    return fail(PGPU_EDEVICE, "initial clear failed");
  }
  *out = d;
  return PGPU_OK;
}

extern "C" int pgpu_domain_destroy(pgpu_domain* d)
{
  if (!d) return PGPU_OK;
  if (d->stream) {
    (void)hipStreamSynchronize(d->stream);
    (void)hipStreamDestroy(d->stream);
  }
  if (d->block) (void)hipFree(d->block);
  if (d->sig) (void)hipFree(d->sig);
  if (d->tape) (void)hipFree(d->tape);
  pgpu_adjoint_free(d);
  delete d->timer;
  delete d;
  return PGPU_OK;
}

extern "C" int pgpu_set_precision(pgpu_domain* d, int precision)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (precision != PGPU_PREC_FP64 && precision != PGPU_PREC_FP32_FLUX &&
      precision != PGPU_PREC_FP64_FAST)
    return fail(PGPU_EINVAL, "unknown precision %d", precision);
  if (precision == PGPU_PREC_FP64_FAST &&
      (d->pipeline != PGPU_PIPELINE_FUSED || d->g.emf_avg != EMF_UCT_HLL ||
       d->g.solver != SOLVER_HLLD || d->g.ghost))
    return fail(PGPU_EINVAL, "the tolerance-mode FP64 flux covers the periodic fused "
                             "HLLD + UCT_HLL step");
  if (precision == PGPU_PREC_FP32_FLUX && d->pipeline != PGPU_PIPELINE_FUSED)
    return fail(PGPU_EINVAL, "the FP32-flux path needs the fused pipeline");
  if (precision == PGPU_PREC_FP32_FLUX && d->g.emf_avg != EMF_UCT_HLL)
    return fail(PGPU_EINVAL, "the FP32-flux path uses the UCT_HLL EMF average");
  if (precision == PGPU_PREC_FP32_FLUX && d->g.ghost)
    return fail(PGPU_EINVAL, "a ghosted box runs on the FP64 pipeline");
  d->precision = precision;
  return PGPU_OK;
}

extern "C" int pgpu_set_emf_average(pgpu_domain* d, int emf_average)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (emf_average != PGPU_EMF_ARITHMETIC && emf_average != PGPU_EMF_UCT0 &&
      emf_average != PGPU_EMF_UCT_CONTACT && emf_average != PGPU_EMF_UCT_HLL)
    return fail(PGPU_EINVAL, "unknown EMF average %d", emf_average);
  if (emf_average != PGPU_EMF_UCT_HLL &&
      (d->precision != PGPU_PREC_FP64 || d->pipeline != PGPU_PIPELINE_FUSED))
    return fail(PGPU_EINVAL, "the ARITHMETIC, UCT0 and UCT_CONTACT averages need the FP64 fused pipeline");
  d->g.emf_avg = emf_average;
  return PGPU_OK;
}

extern "C" int pgpu_set_solver(pgpu_domain* d, int solver)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (solver == PGPU_SOLVER_HLLD) d->g.solver = SOLVER_HLLD;
  else if (solver == PGPU_SOLVER_HLL) d->g.solver = SOLVER_HLL;
  else if (solver == PGPU_SOLVER_HLLC) d->g.solver = SOLVER_HLLC;
  else if (solver == PGPU_SOLVER_TVDLF) d->g.solver = SOLVER_TVDLF;
  else if (solver == PGPU_SOLVER_ROE) d->g.solver = SOLVER_ROE;
  else return fail(PGPU_EINVAL, "unknown Riemann solver %d", solver);
  return PGPU_OK;
}

// A ghosted domain's device index of interior cell (i, j, k) (pgpu::GhostBox,
// uni layout: PLUTO's cell beg + i sits at beg + i + 1).
static long gh_index(const GhostBox& b, int i, int j, int k)
{
  return ((long)(k + b.beg[2] + 1) * (b.t[1] + 1) + (j + b.beg[1] + 1)) * (b.t[0] + 1) +
         (i + b.beg[0] + 1);
}

// Interior array [nz][ny][nx] -> the whole device block of a ghosted domain:
// every cell outside the interior takes the nearest interior value (the
// first stage's ghost fill replaces the ghost zones; the spare planes only
// feed cells outside every reduction).  face_axis >= 0: a face array whose
// entry f along that axis is the face f+1/2; the face below the first cell
// is the image of the last face, as for the periodic domain.
// low (optional): that face plane itself, [k][j], [k][i] or [j][i] for
// face_axis 0, 1, 2.
static long low_index(const GhostBox& b, int a, const int p[3])
{
  return a == 0 ? (long)p[2] * b.n[1] + p[1] : a == 1 ? (long)p[2] * b.n[0] + p[0]
                                                      : (long)p[1] * b.n[0] + p[0];
}

static void gh_expand(const GhostBox& b, const double* in, double* out, int face_axis,
                      const double* low = nullptr)
{
  const int n[3] = {b.n[0], b.n[1], b.n[2]};
  long o = 0;
  for (int K = 0; K <= b.t[2]; K++)
    for (int J = 0; J <= b.t[1]; J++)
      for (int I = 0; I <= b.t[0]; I++, o++) {
        int p[3] = {I - b.beg[0] - 1, J - b.beg[1] - 1, K - b.beg[2] - 1};
        const bool below = face_axis >= 0 && p[face_axis] == -1;
        for (int a = 0; a < 3; a++) {
          if (a == face_axis && p[a] == -1) p[a] = n[a] - 1;
          p[a] = p[a] < 0 ? 0 : (p[a] >= n[a] ? n[a] - 1 : p[a]);
        }
        out[o] = (below && low) ? low[low_index(b, face_axis, p)]
                                : in[((long)p[2] * n[1] + p[1]) * n[0] + p[0]];
      }
}

extern "C" int pgpu_domain_create_bounded(int nx, int ny, int nz, int nghost, double dx, double dy,
                                          double dz, double gamma, const int bound[6],
                                          pgpu_domain** out)
{
  if (!out || !bound) return fail(PGPU_EINVAL, "NULL argument");
  *out = nullptr;
  // PLM + CT reads two cells beyond the interfaces it solves (get_nghost.c
  // gives 2 for LINEAR reconstruction)
  if (nghost < 2 || nx < nghost || ny < nghost || nz < nghost || nx < 3 || ny < 3 || nz < 3)
    return fail(PGPU_EINVAL, "a bounded box needs nghost >= 2 and at least max(3, nghost) cells "
                             "per axis");
  const int t[3] = {nx + 2 * nghost, ny + 2 * nghost, nz + 2 * nghost};
  const int beg[3] = {nghost, nghost, nghost};
  const int end[3] = {nghost + nx - 1, nghost + ny - 1, nghost + nz - 1};
  // Cartesian face areas (set_geometry.c:149,180,202)
  const double area[3] = {1.0 * dy * dz, dx * 1.0 * dz, dx * dy * 1.0};
  pgpu_domain* d = nullptr;
  int rc = pgpu_domain_create(t[0] + 1, t[1] + 1, t[2] + 1, 3, dx, dy, dz, gamma, &d);
  if (rc != PGPU_OK) return rc;
  rc = pgpu_domain_set_ghosted(d, t, beg, end, bound, area);
  if (rc != PGPU_OK) {
    const std::string msg = pgpu_last_error();
    pgpu_domain_destroy(d);
    return fail(rc, "%s", msg.c_str());
  }
  *out = d;
  return PGPU_OK;
}

// set_state of a ghosted domain: the interior cells' arrays, expanded
static int set_state_ghosted(pgpu_domain* d, const double* prim, const double* const faces[3],
                             const double* const* low = nullptr)
{
  const GhostBox& b = d->gb;
  const long ni = (long)b.n[0] * b.n[1] * b.n[2];
  const size_t nb = (size_t)d->g.n * sizeof(double);
  std::vector<double> big((size_t)d->g.n), cb((size_t)ni);
  unsigned int nans[2] = {0, 0};
  std::vector<unsigned char> nan_cell((size_t)ni, 0);
  for (int q = 0; q < 8; q++) {
    const double* src = prim + (size_t)q * ni;
    if (q >= BX1 && q <= BX3) {
      // the cell-centred field: face average with the image face below
      const double* f = faces[q - BX1];
      const int a = q - BX1;
      for (int k = 0; k < b.n[2]; k++)
        for (int j = 0; j < b.n[1]; j++)
          for (int i = 0; i < b.n[0]; i++) {
            const long c = ((long)k * b.n[1] + j) * b.n[0] + i;
            const int p[3] = {i, j, k};
            const long st[3] = {1, b.n[0], (long)b.n[0] * b.n[1]};
            const long cm = c + (p[a] == 0 ? (long)(b.n[a] - 1) * st[a] : -st[a]);
            const double fm = (p[a] == 0 && low) ? low[a][low_index(b, a, p)] : f[cm];
            cb[c] = 0.5 * (f[c] + fm);
          }
      src = cb.data();
    }
    for (long c = 0; c < ni; c++) nan_cell[c] |= (src[c] != src[c]);
    gh_expand(b, src, big.data(), -1);
    HIP_TRY(hipMemcpy(d->V[q], big.data(), nb, hipMemcpyHostToDevice));
  }
  for (int a = 0; a < 3; a++) {
    gh_expand(b, faces[a], big.data(), a, low ? low[a] : nullptr);
    HIP_TRY(hipMemcpy(d->bA[a], big.data(), nb, hipMemcpyHostToDevice));
  }
  for (long c = 0; c < ni; c++) nans[0] += nan_cell[c];
  HIP_TRY(hipMemcpy(&d->sc->nan_cells, nans, sizeof(nans), hipMemcpyHostToDevice));
  return PGPU_OK;
}

int pgpu_ghosted_set_state(pgpu_domain* d, const double* prim, const double* const faces[3],
                           const double* const* low)
{
  if (!d || !d->g.ghost) return fail(PGPU_EINVAL, "not a bounded box");
  HIP_TRY(hipStreamSynchronize(d->stream));
  return set_state_ghosted(d, prim, faces, low);
}

int pgpu_ghosted_low_faces(pgpu_domain* d, double* const* low)
{
  if (!d || !d->g.ghost) return fail(PGPU_EINVAL, "not a bounded box");
  const GhostBox& b = d->gb;
  std::vector<double> big((size_t)d->g.n);
  for (int a = 0; a < 3; a++) {
    HIP_TRY(hipMemcpyAsync(big.data(), d->bA[a], (size_t)d->g.n * sizeof(double),
                           hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    for (int k = 0; k < b.n[2]; k++)
      for (int j = 0; j < b.n[1]; j++)
        for (int i = 0; i < b.n[0]; i++) {
          int p[3] = {i, j, k};
          if (p[a] != 0) continue;
          p[a] = -1;                        // the face below the first cell
          const long dev = gh_index(b, p[0], p[1], p[2]);
          p[a] = 0;
          low[a][low_index(b, a, p)] = big[dev];
        }
  }
  return PGPU_OK;
}

extern "C" int pgpu_set_state_bounded(pgpu_domain* d, const double* prim, const double* bx,
                                      const double* by, const double* bz, const double* bx_low,
                                      const double* by_low, const double* bz_low)
{
  if (!d || !prim || !bx || !by || !bz || !bx_low || !by_low || !bz_low)
    return fail(PGPU_EINVAL, "NULL argument");
  if (!d->g.ghost) return fail(PGPU_EINVAL, "not a bounded box (a periodic box has no lower faces)");
  HIP_TRY(hipStreamSynchronize(d->stream));
  const double* faces[3] = {bx, by, bz};
  const double* low[3] = {bx_low, by_low, bz_low};
  return set_state_ghosted(d, prim, faces, low);
}

extern "C" int pgpu_get_state_bounded(pgpu_domain* d, double* prim, double* bx, double* by,
                                      double* bz, double* bx_low, double* by_low, double* bz_low)
{
  if (!d || !bx_low || !by_low || !bz_low) return fail(PGPU_EINVAL, "NULL argument");
  if (!d->g.ghost) return fail(PGPU_EINVAL, "not a bounded box (a periodic box has no lower faces)");
  const int rc = pgpu_get_state(d, prim, bx, by, bz);
  if (rc != PGPU_OK) return rc;
  double* low[3] = {bx_low, by_low, bz_low};
  return pgpu_ghosted_low_faces(d, low);
}

extern "C" int pgpu_set_state(pgpu_domain* d, const double* prim, const double* bx,
                              const double* by, const double* bz)
{
  if (!d || !prim || !bx || !by || !bz) return fail(PGPU_EINVAL, "NULL argument");
  if (d->g.ghost) {
    HIP_TRY(hipStreamSynchronize(d->stream));
    const double* faces[3] = {bx, by, bz};
    return set_state_ghosted(d, prim, faces);
  }
  const Geo& g = d->g;
  const size_t nb = (size_t)g.n * sizeof(double);
  // cell-centred field = face average, computed on the host in the
  // reference's operand order 0.5*(b_i + b_{i-1})
  std::vector<double> cb((size_t)g.n);
  std::vector<unsigned char> nan_cell((size_t)g.n, 0);
  for (int q = 0; q < 8; q++) {
    if (q >= BX1 && q <= BX3) continue;
    HIP_TRY(hipMemcpyAsync(d->V[q], prim + (size_t)q * g.n, nb, hipMemcpyHostToDevice, d->stream));
    const double* pq = prim + (size_t)q * g.n;
    for (long c = 0; c < g.n; c++) nan_cell[c] |= (pq[c] != pq[c]);
  }
  const double* faces[3] = {bx, by, bz};
  for (int a = 0; a < 3; a++) {
    HIP_TRY(hipMemcpyAsync(d->bA[a], faces[a], nb, hipMemcpyHostToDevice, d->stream));
  }
  for (int a = 0; a < 3; a++) {
    const double* f = faces[a];
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) {
          const long c = ((long)k * g.ny + j) * g.nx + i;
          long cm;
          if (a == 0) cm = c + ((i == 0) ? g.nx - 1 : -1);
          else if (a == 1) cm = c + ((j == 0) ? (long)(g.ny - 1) * g.nx : -(long)g.nx);
          else cm = c + ((k == 0) ? (long)(g.nz - 1) * g.nx * g.ny : -(long)g.nx * g.ny);
          cb[c] = (a == 2 && g.dims == 2) ? 0.0 : 0.5 * (f[c] + f[cm]);
        }
    HIP_TRY(hipMemcpyAsync(d->V[BX1 + a], cb.data(), nb, hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    for (long c = 0; c < g.n; c++) nan_cell[c] |= (cb[c] != cb[c]);
  }
  // the first stage input: cells holding a NaN (CheckNaN, tools.c:20-38);
  // the stages count the ones they produce
  unsigned int nans[2] = {0, 0};   // nan_cells, nan_next
  for (long c = 0; c < g.n; c++) nans[0] += nan_cell[c];
  HIP_TRY(hipMemcpyAsync(&d->sc->nan_cells, nans, sizeof(nans), hipMemcpyHostToDevice, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return PGPU_OK;
}

extern "C" int pgpu_get_state(pgpu_domain* d, double* prim, double* bx, double* by, double* bz)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (d->g.ghost) {
    // the interior cells and their upper faces
    const GhostBox& b = d->gb;
    const long ni = (long)b.n[0] * b.n[1] * b.n[2];
    std::vector<double> big((size_t)d->g.n);
    double* faces[3] = {bx, by, bz};
    for (int q = 0; q < 11; q++) {
      double* dst = q < 8 ? (prim ? prim + (size_t)q * ni : nullptr) : faces[q - 8];
      if (!dst) continue;
      HIP_TRY(hipMemcpyAsync(big.data(), q < 8 ? d->V[q] : d->bA[q - 8],
                             (size_t)d->g.n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
      HIP_TRY(hipStreamSynchronize(d->stream));
      for (int k = 0; k < b.n[2]; k++)
        for (int j = 0; j < b.n[1]; j++)
          for (int i = 0; i < b.n[0]; i++)
            dst[((long)k * b.n[1] + j) * b.n[0] + i] = big[gh_index(b, i, j, k)];
    }
    return PGPU_OK;
  }
  const Geo& g = d->g;
  const size_t nb = (size_t)g.n * sizeof(double);
  if (prim)
    for (int q = 0; q < 8; q++)
      HIP_TRY(hipMemcpyAsync(prim + (size_t)q * g.n, d->V[q], nb, hipMemcpyDeviceToHost, d->stream));
  double* faces[3] = {bx, by, bz};
  for (int a = 0; a < 3; a++)
    if (faces[a])
      HIP_TRY(hipMemcpyAsync(faces[a], d->bA[a], nb, hipMemcpyDeviceToHost, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return PGPU_OK;
}

extern "C" int pgpu_state_device_ptrs(pgpu_domain* d, double** prim8, double** face3)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (d->g.ghost) return fail(PGPU_EINVAL, "a bounded box does not expose its ghosted arrays");
  if (prim8)
    for (int q = 0; q < 8; q++) prim8[q] = d->V[q];
  if (face3)
    for (int a = 0; a < 3; a++) face3[a] = d->bA[a];
  return PGPU_OK;
}

extern "C" int pgpu_set_dt(pgpu_domain* d, double dt, double cfl, double cfl_max_var, int fixed_dt)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (!(dt > 0.0)) return fail(PGPU_EINVAL, "dt must be positive");
  if (!fixed_dt && (!(cfl > 0.0) || !(cfl_max_var > 0.0)))
    return fail(PGPU_EINVAL, "cfl and cfl_max_var must be positive");
  d->cfl = cfl;
  d->cfl_max_var = cfl_max_var;
  d->fixed_dt = fixed_dt ? 1 : 0;
  DevScalars h;
  memset(&h, 0, sizeof(h));
  h.dt = dt;
  h.first_dt = dt;
  HIP_TRY(hipMemcpyAsync(d->sc, &h, sizeof(h), hipMemcpyHostToDevice, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return PGPU_OK;
}

static int enqueue_step_modular(pgpu_domain* d)
{
  const Geo& g = d->g;
  KernelTimer* kt = d->timer;
  StageIn in1, in2, vn;
  for (int q = 0; q < 8; q++) {
    in1.V[q] = d->V[q];
    in2.V[q] = d->V1[q];
    vn.V[q] = d->V[q];
  }
  for (int a = 0; a < 3; a++) {
    in1.b[a] = d->bA[a];
    in2.b[a] = d->bB[a];
    vn.b[a] = d->bA[a];
  }
  Cons U;
  for (int q = 0; q < 5; q++) U.u[q] = d->U[q];
  Prims V1, V;
  for (int q = 0; q < 8; q++) {
    V1.v[q] = d->V1[q];
    V.v[q] = d->V[q];
  }
  FaceIn fA, fB;
  FaceOut oA, oB;
  for (int a = 0; a < 3; a++) {
    fA.b[a] = d->bA[a];
    fB.b[a] = d->bB[a];
    oA.b[a] = d->bA[a];
    oB.b[a] = d->bB[a];
  }
  int rc;
  if (d->tape_mode) {
    // branch tape: the stages record or replay their piecewise outcomes
    const int replay = d->tape_mode == 2;
    if (d->tape_step >= (replay ? d->tape_len : d->tape_cap))
      return fail(PGPU_EINVAL, "step %d is past the branch tape (%d steps)", d->tape_step,
                  replay ? d->tape_len : d->tape_cap);
    const size_t per_stage = 4 * (size_t)g.n;
    unsigned long long* t1 = d->tape + (size_t)(2 * d->tape_step) * per_stage;
    rc = launch_stage_taped(g, 1, in1, d->sw, U, d->e, fA, fA, oB, vn, V1, d->flag, d->sc, t1,
                            replay, d->stream);
    if (!rc)
      rc = launch_stage_taped(g, 2, in2, d->sw, U, d->e, fB, fA, oA, vn, V, d->flag, d->sc,
                              t1 + per_stage, replay, d->stream);
    if (rc) return fail(PGPU_EDEVICE, "taped stage launch failed: %s", hipGetErrorString((hipError_t)rc));
    d->tape_step++;
    if (!replay && d->tape_step > d->tape_len) d->tape_len = d->tape_step;
    return PGPU_OK;
  }
  // stage 1: V, bA -> U*, bB, V1
  rc = launch_stage(g, 1, in1, d->sw, U, d->e, fA, fA, oB, vn, V1, d->flag, d->sc, d->stream, kt);
  if (rc) return fail(PGPU_EDEVICE, "stage-1 launch failed: %s", hipGetErrorString((hipError_t)rc));
  // stage 2: V1, bB (+ V, bA for the blend) -> V, bA
  rc = launch_stage(g, 2, in2, d->sw, U, d->e, fB, fA, oA, vn, V, d->flag, d->sc, d->stream, kt);
  if (rc) return fail(PGPU_EDEVICE, "stage-2 launch failed: %s", hipGetErrorString((hipError_t)rc));
  return PGPU_OK;
}

// Boundary() on a ghosted domain (boundary.c:120-292): the ghost zones of
// the stage input (prims v, faces b), side by side.
// Copies the slab of a USERDEF side between the device box and the host
// staging (full device extents along the two transverse axes, the slab's
// normal range along the side's axis) for all 11 arrays.
static int slab_copy(pgpu_domain* d, const pgpu_slab& s, int axis, double* const* v,
                     double* const* b, bool to_host)
{
  const GhostBox& g = d->gb;
  const size_t T0 = (size_t)g.t[0] + 1, T1 = (size_t)g.t[1] + 1;
  const size_t per = (size_t)s.ld[0] * s.ld[1] * s.ld[2];
  const int l = s.lo[axis] + 1;   // device index of the first slab plane
  size_t width, height, pitch, start;
  if (axis == 0) {
    width = (size_t)s.ld[0] * sizeof(double);
    height = T1 * (g.t[2] + 1);
    pitch = T0 * sizeof(double);
    start = (size_t)l;
  } else if (axis == 1) {
    width = (size_t)s.ld[1] * T0 * sizeof(double);
    height = (size_t)g.t[2] + 1;
    pitch = T1 * T0 * sizeof(double);
    start = (size_t)l * T0;
  } else {
    width = (size_t)s.ld[2] * T1 * T0 * sizeof(double);
    height = 1;
    pitch = width;
    start = (size_t)l * T1 * T0;
  }
  for (int q = 0; q < 11; q++) {
    double* dev = (q < 8 ? v[q] : b[q - 8]) + start;
    double* host = d->slab_buf.data() + (size_t)q * per;
    const hipError_t e =
        to_host ? hipMemcpy2DAsync(host, width, dev, pitch, width, height, hipMemcpyDeviceToHost,
                                   d->stream)
                : hipMemcpy2DAsync(dev, pitch, host, width, width, height, hipMemcpyHostToDevice,
                                   d->stream);
    if (e != hipSuccess) return (int)e;
  }
  return (int)hipStreamSynchronize(d->stream);
}

// A USERDEF side (boundary.c:258-292): the stage input's slab goes to the
// host hook and back, then FillMagneticField and CT_AverageNormalMagField
// run on the device.
static int userdef_side(pgpu_domain* d, int is, int stage, double* const* v, double* const* b)
{
  if (!d->userdef)
    return fail(PGPU_EINVAL, "side %d is user-defined (USERDEF) but no hook is set "
                             "(pgpu_set_userdef_boundary)", is);
  const GhostBox& g = d->gb;
  const int a = is / 2;
  pgpu_slab s;
  for (int x = 0; x < 3; x++) {
    s.lo[x] = -1;
    s.hi[x] = g.t[x] - 1;
    s.beg[x] = g.beg[x];
    s.end[x] = g.end[x];
    s.tot[x] = g.t[x];
  }
  if (!(is & 1)) s.hi[a] = 2 * g.beg[a] - 1;                  // ghosts + as many interior layers
  else s.lo[a] = g.end[a] - (g.t[a] - 1 - g.end[a]);           // (and the face below them)
  for (int x = 0; x < 3; x++) s.ld[x] = s.hi[x] - s.lo[x] + 1;
  const size_t per = (size_t)s.ld[0] * s.ld[1] * s.ld[2];
  d->slab_buf.resize(11 * per);
  for (int q = 0; q < 8; q++) s.vc[q] = d->slab_buf.data() + (size_t)q * per;
  for (int c = 0; c < 3; c++) s.vs[c] = d->slab_buf.data() + (size_t)(8 + c) * per;
  int rc = slab_copy(d, s, a, v, b, true);
  if (rc) return fail(PGPU_EDEVICE, "USERDEF side %d: slab download failed (%d)", is, rc);
  if (d->userdef(d->userdef_ctx, is, stage, &s) != 0)
    return fail(PGPU_EINVAL, "USERDEF side %d: the boundary hook failed", is);
  rc = slab_copy(d, s, a, v, b, false);
  if (rc) return fail(PGPU_EDEVICE, "USERDEF side %d: slab upload failed (%d)", is, rc);
  rc = launch_boundary_fill_normal(g, is, v, b, nullptr, d->area, d->stream);
  if (rc) return fail(PGPU_EDEVICE, "USERDEF side %d: fill launch failed (%d)", is, rc);
  return PGPU_OK;
}

static int enqueue_boundary(pgpu_domain* d, double* const* v, double* const* b, int stage)
{
  if (!d->g.ghost) return 0;
  KernelTimer* kt = d->timer;
  if (kt) kt->begin(KT_BOUNDARY, d->stream);
  int rc = 0;
  for (int is = 0; is < 6 && !rc; is++) {
    if (d->bound[is] == BOUND_USERDEF) {
      rc = userdef_side(d, is, stage, v, b);
      if (rc) {
        if (kt) kt->end(KT_BOUNDARY, d->stream);
        return rc;
      }
      continue;
    }
    rc = launch_boundary_side(d->gb, is, d->bound[is], v, b, nullptr, d->area, d->stream);
  }
  if (kt) kt->end(KT_BOUNDARY, d->stream);
  if (rc) return fail(PGPU_EDEVICE, "boundary launch failed: %s", hipGetErrorString((hipError_t)rc));
  return PGPU_OK;
}

extern "C" int pgpu_set_branch_signature(pgpu_domain* d, int on)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (!on) {
    d->sig_on = false;
    return PGPU_OK;
  }
  if (d->pipeline != PGPU_PIPELINE_FUSED || d->precision != PGPU_PREC_FP64 ||
      d->g.solver != SOLVER_HLLD || d->g.emf_avg != EMF_UCT_HLL || d->g.ghost)
    return fail(PGPU_EINVAL, "the branch signature covers the periodic fused FP64 HLLD + "
                             "UCT_HLL step");
  if (d->tape_mode) return fail(PGPU_EINVAL, "the branch signature and the branch tape exclude each other");
  if (!d->sig) HIP_TRY(hipMalloc(&d->sig, sizeof(unsigned long long)));
  HIP_TRY(hipStreamSynchronize(d->stream));
  HIP_TRY(hipMemset(d->sig, 0, sizeof(unsigned long long)));
  d->sig_on = true;
  return PGPU_OK;
}

extern "C" int pgpu_set_branch_tape(pgpu_domain* d, int mode, int max_steps)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (mode == PGPU_TAPE_OFF) {
    HIP_TRY(hipStreamSynchronize(d->stream));
    if (d->tape) (void)hipFree(d->tape);
    d->tape = nullptr;
    d->tape_mode = d->tape_cap = d->tape_step = d->tape_len = 0;
    return PGPU_OK;
  }
  if (mode != PGPU_TAPE_RECORD && mode != PGPU_TAPE_REPLAY)
    return fail(PGPU_EINVAL, "unknown branch tape mode %d", mode);
  if (d->precision != PGPU_PREC_FP64 || d->g.solver != SOLVER_HLLD ||
      d->g.emf_avg != EMF_UCT_HLL || d->g.ghost)
    return fail(PGPU_EINVAL, "the branch tape covers the periodic FP64 HLLD + UCT_HLL step");
  if (d->sig_on) return fail(PGPU_EINVAL, "the branch signature and the branch tape exclude each other");
  if (mode == PGPU_TAPE_REPLAY) {
    if (!d->tape_len) return fail(PGPU_EINVAL, "no branch tape was recorded");
    d->tape_mode = mode;
    d->tape_step = 0;
    return PGPU_OK;
  }
  if (max_steps < 1) return fail(PGPU_EINVAL, "max_steps must be >= 1");
  const size_t bytes = (size_t)max_steps * 2 * 4 * (size_t)d->g.n * sizeof(unsigned long long);
  HIP_TRY(hipStreamSynchronize(d->stream));
  if (max_steps > d->tape_cap) {
    if (d->tape) (void)hipFree(d->tape);
    d->tape = nullptr;
    d->tape_cap = 0;
    HIP_TRY(hipMalloc(&d->tape, bytes));
    d->tape_cap = max_steps;
  }
  d->tape_mode = mode;
  d->tape_step = 0;
  d->tape_len = 0;
  return PGPU_OK;
}

extern "C" int pgpu_get_branch_signature(pgpu_domain* d, unsigned long long* out)
{
  if (!d || !out) return fail(PGPU_EINVAL, "NULL argument");
  if (!d->sig) return fail(PGPU_EINVAL, "no branch signature was recorded");
  unsigned long long h = 0;
  unsigned int fails = 0;
  HIP_TRY(hipMemcpyAsync(&h, d->sig, sizeof(h), hipMemcpyDeviceToHost, d->stream));
  HIP_TRY(hipMemcpyAsync(&fails, &d->sc->fail_cells, sizeof(fails), hipMemcpyDeviceToHost,
                         d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  // ConsToPrim floors (the only other branch of the step) count too
  *out = h + 0x9E3779B97F4A7C15ull * (unsigned long long)fails;
  return PGPU_OK;
}

extern "C" int pgpu_set_userdef_boundary(pgpu_domain* d, pgpu_userdef_fn fn, void* ctx)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  d->userdef = fn;
  d->userdef_ctx = ctx;
  return PGPU_OK;
}

static int enqueue_step_fused(pgpu_domain* d)
{
  const Geo& g = d->g;
  KernelTimer* kt = d->timer;
  const bool fp32 = d->precision == PGPU_PREC_FP32_FLUX;
  const bool fast = d->precision == PGPU_PREC_FP64_FAST;
  if (fast && (d->g.solver != SOLVER_HLLD || d->g.emf_avg != EMF_UCT_HLL))
    return fail(PGPU_EINVAL, "the tolerance-mode FP64 flux covers HLLD + UCT_HLL");
  auto flux = [&](int stage, const StageIn& in, const Cons& U, const FaceData& fd) {
    return fast ? launch_flux_fused_fast(g, stage, in, U, fd, d->cdt, d->sc, d->stream, kt)
                : launch_flux_fused(g, stage, fp32, in, U, fd, d->cdt, d->sc, d->stream, kt);
  };
  StageIn in1, in2, vn;
  for (int q = 0; q < 8; q++) {
    in1.V[q] = d->V[q];
    in2.V[q] = d->V1[q];
    vn.V[q] = d->V[q];
  }
  for (int a = 0; a < 3; a++) {
    in1.b[a] = d->bA[a];
    in2.b[a] = d->bB[a];
    vn.b[a] = d->bA[a];
  }
  Cons U;
  for (int q = 0; q < 5; q++) U.u[q] = d->U[q];
  Prims V1, V;
  for (int q = 0; q < 8; q++) {
    V1.v[q] = d->V1[q];
    V.v[q] = d->V[q];
  }
  FaceData fd;
  for (int a = 0; a < 3; a++) {
    fd.SL[a] = d->sw.d[a].SL;
    fd.SR[a] = d->sw.d[a].SR;
    for (int c = 0; c < 3; c++) fd.dv[a][c] = d->sw.d[a].dv[c];
  }
  FaceIn fA, fB;
  FaceOut oB, oC;
  for (int a = 0; a < 3; a++) {
    fA.b[a] = d->bA[a];
    fB.b[a] = d->bB[a];
    oB.b[a] = d->bB[a];
    oC.b[a] = d->bC[a];
  }
  // stage 1: V, bA -> U*, bB, V1
  int rc = enqueue_boundary(d, d->V, d->bA, 1);
  if (rc) return rc;
  const bool sig = d->sig_on;
  if (sig) rc = launch_sig_sweep(g, in1, d->sig, d->stream);
  if (!rc) rc = flux(1, in1, U, fd);
  if (!rc && sig) rc = launch_sig_emf(g, in1, d->sw, d->sig, d->stream);
  if (!rc) rc = launch_emf(g, in1, d->sw, d->e, d->stream, kt, fp32);
  // the FP32-flux path stores V1 (read only by stage 2's FP32 flux and EMF
  // kernels) as float
  if (!rc) rc = launch_ct_c2p(g, 1, d->e, fA, fA, oB, U, vn, V1, d->flag, d->sc, d->stream, kt, fp32);
  // stage 2: V1, bB (+ V, bA for the blend) -> V, bC
  if (rc) return fail(PGPU_EDEVICE, "stage launch failed: %s", hipGetErrorString((hipError_t)rc));
  rc = enqueue_boundary(d, d->V1, d->bB, 2);
  if (rc) return rc;
  if (sig) rc = launch_sig_sweep(g, in2, d->sig, d->stream);
  if (!rc) rc = flux(2, in2, U, fd);
  if (!rc && sig) rc = launch_sig_emf(g, in2, d->sw, d->sig, d->stream);
  if (!rc) rc = launch_emf(g, in2, d->sw, d->e, d->stream, kt, fp32, fp32);
  if (!rc) rc = launch_ct_c2p(g, 2, d->e, fB, fA, oC, U, vn, V, d->flag, d->sc, d->stream, kt);
  if (rc) return fail(PGPU_EDEVICE, "stage launch failed: %s", hipGetErrorString((hipError_t)rc));
  // the step's faces are in bC: rotate so that bA is the state again
  for (int a = 0; a < 3; a++) {
    double* t = d->bA[a];
    d->bA[a] = d->bC[a];
    d->bC[a] = t;
  }
  return PGPU_OK;
}

static int enqueue_step(pgpu_domain* d)
{
  int rc = launch_zero_flags(d->g, d->flag, d->stream, d->sc);
  if (rc) return fail(PGPU_EDEVICE, "flag clear launch failed (%d)", rc);
  rc = (d->pipeline == PGPU_PIPELINE_MODULAR || d->tape_mode) ? enqueue_step_modular(d)
                                                              : enqueue_step_fused(d);
  if (rc) return rc;
  KernelTimer* kt = d->timer;
  if (kt) kt->begin(KT_STEP_END, d->stream);
  rc = launch_step_end(d->sc, d->g.dims, d->cfl, d->cfl_max_var, d->fixed_dt, d->stream);
  if (kt) kt->end(KT_STEP_END, d->stream);
  if (rc) return fail(PGPU_EDEVICE, "step-end launch failed (%d)", rc);
  return PGPU_OK;
}

int pgpu_domain_set_ghosted(pgpu_domain* d, const int t[3], const int beg[3], const int end[3],
                            const int type[6], const double area[3])
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (d->g.dims != 3 || d->g.nx != t[0] + 1 || d->g.ny != t[1] + 1 || d->g.nz != t[2] + 1)
    return fail(PGPU_EINVAL, "a ghosted box needs a 3D domain of (NX_TOT + 1) cells per axis");
  if (d->pipeline != PGPU_PIPELINE_FUSED || d->precision != PGPU_PREC_FP64)
    return fail(PGPU_EINVAL, "a ghosted box runs on the fused FP64 pipeline");
  for (int is = 0; is < 6; is++)
    if ((type[is] < BOUND_OUTFLOW || type[is] > BOUND_PERIODIC) && type[is] != BOUND_USERDEF)
      return fail(PGPU_EINVAL, "boundary type %d is not supported on the device", type[is]);
  for (int a = 0; a < 3; a++)
    if ((type[2 * a] == BOUND_PERIODIC) != (type[2 * a + 1] == BOUND_PERIODIC))
      return fail(PGPU_EINVAL, "axis %d has only one periodic side", a);
  GhostBox gb;
  for (int a = 0; a < 3; a++) {
    const int n = end[a] - beg[a] + 1;
    if (beg[a] < 2 || end[a] > t[a] - 3 || n < beg[a] || n < t[a] - 1 - end[a])
      return fail(PGPU_EINVAL, "unexpected ghosted extents along axis %d", a);
    gb.t[a] = t[a];
    gb.beg[a] = beg[a];
    gb.end[a] = end[a];
    gb.n[a] = n;
    d->g.lo[a] = beg[a] + 1;
    d->g.hi[a] = end[a] + 1;
    d->area[a] = area[a];
  }
  gb.uni = 1;
  d->gb = gb;
  for (int is = 0; is < 6; is++) d->bound[is] = type[is];
  d->g.ghost = 1;
  return PGPU_OK;
}

extern "C" int pgpu_set_pipeline(pgpu_domain* d, int pipeline)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (d->g.ghost && pipeline != PGPU_PIPELINE_FUSED)
    return fail(PGPU_EINVAL, "a ghosted box runs on the fused pipeline");
  if (pipeline != PGPU_PIPELINE_FUSED && pipeline != PGPU_PIPELINE_MODULAR)
    return fail(PGPU_EINVAL, "unknown pipeline %d", pipeline);
  if (pipeline != PGPU_PIPELINE_FUSED && d->precision != PGPU_PREC_FP64)
    return fail(PGPU_EINVAL, "the modular pipeline is FP64 only");
  if (pipeline != PGPU_PIPELINE_FUSED && d->g.emf_avg != EMF_UCT_HLL)
    return fail(PGPU_EINVAL, "the modular pipeline uses the UCT_HLL EMF average");
  d->pipeline = pipeline;
  return PGPU_OK;
}

extern "C" int pgpu_set_emf_tiling(int bx, int by, int kz)
{
  if (set_emf_tiling(bx, by, kz))
    return fail(PGPU_EINVAL, "EMF tiling %d x %d x %d: bx*by must be a multiple of 64 and <= 256, kz >= 1",
                bx, by, kz);
  return PGPU_OK;
}

extern "C" int pgpu_set_emf_kernel(int kind, int kz)
{
  if (set_emf_kernel(kind, kz)) return fail(PGPU_EINVAL, "EMF kernel %d (kz %d): kind 0, 1 or 2, kz >= 0", kind, kz);
  return PGPU_OK;
}

extern "C" int pgpu_set_flux_kernel(int kind, int seg)
{
  if (set_flux_xy_kernel(kind, seg))
    return fail(PGPU_EINVAL, "flux kernel %d (seg %d): kind 0 or 1, seg >= 0", kind, seg);
  return PGPU_OK;
}

extern "C" int pgpu_get_emf_tiling(int* bx, int* by, int* kz)
{
  if (!bx || !by || !kz) return fail(PGPU_EINVAL, "NULL output");
  get_emf_tiling(bx, by, kz);
  return PGPU_OK;
}

extern "C" int pgpu_run(pgpu_domain* d, int nsteps)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (nsteps < 0) return fail(PGPU_EINVAL, "nsteps must be >= 0");
  for (int s = 0; s < nsteps; s++) {
    int rc = enqueue_step(d);
    if (rc) return rc;
  }
  return PGPU_OK;
}

extern "C" int pgpu_sync(pgpu_domain* d)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  unsigned int nans = 0;
  HIP_TRY(hipMemcpyAsync(&nans, &d->sc->nan_cells, sizeof(nans), hipMemcpyDeviceToHost, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  if (nans)
    return fail(PGPU_ENAN, "NaN found in a stage input (%u cells so far); PLUTO's CheckNaN "
                           "stops the run there", nans);
  unsigned int small = 0;
  double small_dt = 0.0;
  HIP_TRY(hipMemcpyAsync(&small, &d->sc->dt_small, sizeof(small), hipMemcpyDeviceToHost,
                         d->stream));
  HIP_TRY(hipMemcpyAsync(&small_dt, &d->sc->dt_small_value, sizeof(small_dt),
                         hipMemcpyDeviceToHost, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  if (small)
    return fail(PGPU_EDTSMALL, "dt is too small (%12.6e < 1e-9 first_dt); PLUTO's NextTimeStep "
                               "stops the run there, and the steps after it took dt = 0",
                small_dt);
  return PGPU_OK;
}

extern "C" int pgpu_get_scalars(pgpu_domain* d, double out[6])
{
  if (!d || !out) return fail(PGPU_EINVAL, "NULL argument");
  DevScalars h;
  HIP_TRY(hipMemcpyAsync(&h, d->sc, offsetof(DevScalars, dt_hist), hipMemcpyDeviceToHost,
                         d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  out[0] = h.dt;
  out[1] = h.invdt_last;
  out[2] = h.mach_last;
  out[3] = h.time;
  out[4] = (double)h.step;
  out[5] = (double)h.fail_cells;
  return PGPU_OK;
}

extern "C" int pgpu_get_history(pgpu_domain* d, int n, double* dt, double* invdt, double* mach)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  if (n < 0 || n > DT_HISTORY) return fail(PGPU_EINVAL, "n must be in [0, %d]", DT_HISTORY);
  const size_t nb = (size_t)n * sizeof(double);
  if (dt) HIP_TRY(hipMemcpyAsync(dt, d->sc->dt_hist, nb, hipMemcpyDeviceToHost, d->stream));
  if (invdt) HIP_TRY(hipMemcpyAsync(invdt, d->sc->invdt_hist, nb, hipMemcpyDeviceToHost, d->stream));
  if (mach) HIP_TRY(hipMemcpyAsync(mach, d->sc->mach_hist, nb, hipMemcpyDeviceToHost, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return PGPU_OK;
}

extern "C" int pgpu_get_flags(pgpu_domain* d, unsigned char* flag)
{
  if (!d || !flag) return fail(PGPU_EINVAL, "NULL argument");
  if (d->g.ghost) {
    const GhostBox& b = d->gb;
    std::vector<unsigned char> big((size_t)d->g.n);
    HIP_TRY(hipMemcpyAsync(big.data(), d->flag, (size_t)d->g.n, hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    for (int k = 0; k < b.n[2]; k++)
      for (int j = 0; j < b.n[1]; j++)
        for (int i = 0; i < b.n[0]; i++)
          flag[((long)k * b.n[1] + j) * b.n[0] + i] = big[gh_index(b, i, j, k)];
    return PGPU_OK;
  }
  HIP_TRY(hipMemcpyAsync(flag, d->flag, (size_t)d->g.n, hipMemcpyDeviceToHost, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return PGPU_OK;
}

extern "C" int pgpu_timing_enable(pgpu_domain* d, int enable)
{
  if (!d) return fail(PGPU_EINVAL, "domain is NULL");
  HIP_TRY(hipStreamSynchronize(d->stream));
  delete d->timer;
  d->timer = nullptr;
  if (enable) {
    d->timer = new (std::nothrow) EventTimer();
    if (!d->timer) return fail(PGPU_ENOMEM, "timer allocation failed");
  }
  return PGPU_OK;
}

extern "C" int pgpu_timing_read(pgpu_domain* d, int nid, double* ms, long* count)
{
  if (!d || !d->timer) return fail(PGPU_EINVAL, "timing is not enabled");
  HIP_TRY(hipStreamSynchronize(d->stream));
  d->timer->collect();
  for (int q = 0; q < nid && q < KT_COUNT; q++) {
    if (ms) ms[q] = d->timer->ms[q];
    if (count) count[q] = d->timer->count[q];
  }
  return PGPU_OK;
}

extern "C" int pgpu_time_steps(pgpu_domain* d, int nsteps, double* ms)
{
  if (!d || !ms) return fail(PGPU_EINVAL, "NULL argument");
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  HIP_TRY(hipEventRecord(a, d->stream));
  int rc = pgpu_run(d, nsteps);
  HIP_TRY(hipEventRecord(b, d->stream));
  HIP_TRY(hipEventSynchronize(b));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = t;
  return rc;
}

extern "C" int pgpu_domain_bytes(pgpu_domain* d, size_t* bytes)
{
  if (!d || !bytes) return fail(PGPU_EINVAL, "NULL argument");
  *bytes = d->bytes;
  return PGPU_OK;
}
