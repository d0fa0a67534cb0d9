// PLUTO's own entry points for the unsplit MHD stage (include/pluto_abi.h),
// implemented on the device.  Host arrays keep PLUTO's layout (Vc SoA with
// ghost zones, Uc AoS, Vs ArrayBox with index -1, row pointers); the
// interior is exchanged with a device-resident domain per call.
//
//   AdvanceStep              Time_Stepping/rk_step.c:27-254
//   UpdateStage              Time_Stepping/update_stage.c:37-314
//   HLLD_Solver              MHD/hlld.c:44-452
//   HLL_Solver               MHD/hll.c:30-169
//   HLLC_Solver              MHD/hllc.c:32-240
//   LF_Solver                MHD/tvdlf.c:37-163
//   Roe_Solver               MHD/roe.c:53-700
//   States                   States/plm_states.c:80-312
//   PrimToCons, ConsToPrim   MHD/mappers.c:25-254
//   CT_AverageMagneticField  MHD/CT/ct_field_average.c:13-132
//   SetSolver                MHD/set_solver.c:13-81
//   CT_*                     MHD/CT/ct.h:94-112
//   Boundary                 boundary.c:40-315

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "../../include/pluto_abi.h"
#include "boundary_kernels.hpp"
#include "domain.hpp"
#include "mhd_device.hpp"
#include "row_kernels.hpp"

using namespace pgpu;

// The problem file's hook for USERDEF sides: a weak reference, null in a
// program that does not define it.
extern "C" void UserDefBoundary(const Data* d, RBox* box, int side, Grid* grid)
    __attribute__((weak));

// ---------------------------------------------------------------------
// The host program's globals (globals.h).  Weak: a PLUTO executable's own
// definitions take precedence at link and load time.
// ---------------------------------------------------------------------
extern "C" {
__attribute__((weak)) int g_dir = 0;
__attribute__((weak)) int g_intStage = 1;
__attribute__((weak)) double g_dt = 0.0;
__attribute__((weak)) double g_gamma = 5.0 / 3.0;
__attribute__((weak)) double g_maxMach = 0.0;
__attribute__((weak)) int g_i = 0;
__attribute__((weak)) int g_j = 0;
__attribute__((weak)) int g_k = 0;
__attribute__((weak)) long NX1_TOT = 0;
__attribute__((weak)) long NX2_TOT = 0;
__attribute__((weak)) long NX3_TOT = 0;
}

namespace {

constexpr int kPeriodic = 5;   // PERIODIC in pluto.h

void report(const char* fn, const char* what)
{
  fprintf(stderr, "! %s (pluto_amd): %s\n", fn, what);
}

// Device scratch and stream for the row-level entry points.
struct Scratch {
  hipStream_t s = nullptr;
  void* dev = nullptr;
  size_t cap = 0;
  bool init()
  {
    if (s) return true;
    return hipStreamCreate(&s) == hipSuccess;
  }
  double* get(size_t bytes)
  {
    if (!init()) return nullptr;
    if (bytes > cap) {
      if (dev) (void)hipFree(dev);
      dev = nullptr;
      cap = 0;
      if (hipMalloc(&dev, bytes) != hipSuccess) return nullptr;
      cap = bytes;
    }
    return (double*)dev;
  }
};

Scratch& scratch()
{
  static Scratch s;
  return s;
}

// Pinned host staging of AdvanceStep on a periodic box: the 8 cell arrays
// and 3 face arrays of the box, the flags and the device scalars.  Each
// array is packed and its copy enqueued before the next is packed, and
// each download has an event so an array is unpacked while the next is
// still copying.  Timing events bracket the uploads, the step and the
// downloads.
struct Staging {
  double* h = nullptr;           // 11 arrays of n doubles
  unsigned char* fl = nullptr;   // n flags
  DevScalars* sc = nullptr;      // the device scalars after the step
  size_t n = 0;
  hipEvent_t ev[13] = {};        // downloads of the 11 arrays, the flags, the scalars
  hipEvent_t t[4] = {};          // uploads start / end (= step start), step end, downloads end
  bool get(size_t cells)
  {
    if (!t[0]) {
      for (auto& e : ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return false;
      for (auto& e : t)
        if (hipEventCreate(&e) != hipSuccess) return false;
    }
    if (cells <= n) return true;
    release_buffers();
    if (hipHostMalloc((void**)&h, 11 * cells * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&fl, cells, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&sc, sizeof(DevScalars), hipHostMallocDefault) != hipSuccess) {
      release_buffers();
      return false;
    }
    n = cells;
    return true;
  }
  void release_buffers()
  {
    if (h) (void)hipHostFree(h);
    if (fl) (void)hipHostFree(fl);
    if (sc) (void)hipHostFree(sc);
    h = nullptr;
    fl = nullptr;
    sc = nullptr;
    n = 0;
  }
  void release()
  {
    release_buffers();
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : t)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : ev) e = nullptr;
    for (auto& e : t) e = nullptr;
  }
};

Staging& staging()
{
  static Staging s;
  return s;
}

pgpu_dropin_times g_times = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};

void release_scratch()
{
  Scratch& s = scratch();
  if (s.dev) (void)hipFree(s.dev);
  if (s.s) (void)hipStreamDestroy(s.s);
  s.dev = nullptr;
  s.s = nullptr;
  s.cap = 0;
  staging().release();
}

// Geometry of a PLUTO box, checked for what the device path supports.
struct BoxInfo {
  int n[3], beg[3], end[3], dims;
  double dx[3];
  // non-periodic sides (AdvanceStep only): the box is held with its ghosts
  bool ghosted = false;
  int t[3] = {0, 0, 0}, type[6] = {0, 0, 0, 0, 0, 0};
  double area[3] = {0.0, 0.0, 0.0};
};

bool box_of(const Grid* grid, BoxInfo& b, const char* fn, bool allow_ghosted = false)
{
  b.dims = (grid->np_int[2] > 1) ? 3 : 2;
  if (b.dims != 3) {
    // a 2D build with COMPONENTS 3 evolves a cell-centred Bz through its
    // flux, which the face-based device path does not carry
    report(fn, "the PLUTO entry points support DIMENSIONS 3 (COMPONENTS 3) builds");
    return false;
  }
  for (int a = 0; a < 3; a++) {
    if (a == 2 && b.dims == 2) {
      b.n[a] = 1;
      b.beg[a] = b.end[a] = grid->lbeg[a] > 0 ? grid->lbeg[a] : 0;
      b.dx[a] = 1.0;
      continue;
    }
    b.n[a] = grid->np_int[a];
    b.beg[a] = grid->lbeg[a];
    b.end[a] = grid->lend[a];
    b.t[a] = grid->np_tot[a];
    b.type[2 * a] = grid->lbound[a];
    b.type[2 * a + 1] = grid->rbound[a];
    for (int s = 2 * a; s < 2 * a + 2; s++) {
      if (b.type[s] == kPeriodic) continue;
      if (allow_ghosted && b.type[s] == BOUND_USERDEF && !UserDefBoundary) {
        report(fn, "a USERDEF side needs the program's UserDefBoundary");
        return false;
      }
      if (!allow_ghosted ||
          ((b.type[s] < BOUND_OUTFLOW || b.type[s] > BOUND_EQTSYMMETRIC) &&
           b.type[s] != BOUND_USERDEF)) {
        report(fn, allow_ghosted ? "outflow, reflective, axisymmetric, eqtsymmetric, userdef and "
                                   "periodic boundaries are supported on the device"
                                 : "only periodic boundaries are supported on the device");
        return false;
      }
      b.ghosted = true;
    }
    if (grid->nproc[a] > 1) {
      report(fn, "a domain split over processes is not supported");
      return false;
    }
    if (b.end[a] - b.beg[a] + 1 != b.n[a] || b.n[a] < 3) {
      report(fn, "unexpected grid extents");
      return false;
    }
    const double* dx = grid->dx[a];
    b.dx[a] = dx[b.beg[a]];
    for (int i = b.beg[a]; i <= b.end[a]; i++)
      if (dx[i] != b.dx[a]) {
        report(fn, "only uniform grids are supported on the device");
        return false;
      }
  }
  return true;
}

// One cached domain for the whole-box entry points (released by
// pgpu_finalize, or replaced when the grid changes).
pgpu_domain* g_box_dom = nullptr;

pgpu_domain* domain_for(const BoxInfo& b, const char* fn)
{
  pgpu_domain*& dom = g_box_dom;
  static BoxInfo key;
  static double key_gamma = 0.0;
  bool same = dom && key_gamma == g_gamma && key.dims == b.dims && key.ghosted == b.ghosted;
  for (int a = 0; same && a < 3; a++) same = key.n[a] == b.n[a] && key.dx[a] == b.dx[a];
  for (int a = 0; same && b.ghosted && a < 3; a++)
    same = key.t[a] == b.t[a] && key.beg[a] == b.beg[a] && key.area[a] == b.area[a];
  for (int s = 0; same && b.ghosted && s < 6; s++) same = key.type[s] == b.type[s];
  if (same) return dom;
  if (dom) pgpu_domain_destroy(dom);
  dom = nullptr;
  // a ghosted box is held whole, one spare plane below along each axis
  const int m[3] = {b.ghosted ? b.t[0] + 1 : b.n[0], b.ghosted ? b.t[1] + 1 : b.n[1],
                    b.ghosted ? b.t[2] + 1 : b.n[2]};
  if (pgpu_domain_create(m[0], m[1], m[2], b.dims, b.dx[0], b.dx[1], b.dx[2], g_gamma, &dom) !=
          PGPU_OK ||
      (b.ghosted && pgpu_domain_set_ghosted(dom, b.t, b.beg, b.end, b.type, b.area) != PGPU_OK)) {
    report(fn, pgpu_last_error());
    if (dom) pgpu_domain_destroy(dom);
    dom = nullptr;
    return nullptr;
  }
  key = b;
  key_gamma = g_gamma;
  return dom;
}

// A ghosted box on the device (pgpu::GhostBox with uni = 1): PLUTO's
// element (k, j, i) at (k+1, j+1, i+1) of a (t+1)^3 block.  The spare
// plane -1 of an axis where PLUTO's array has none repeats plane 0.
void pack_ghosted(double* h, double*** a, const int t[3], int ext)
{
  const long n0 = t[0] + 1, n1 = t[1] + 1;
  for (int k = -1; k < t[2]; k++)
    for (int j = -1; j < t[1]; j++) {
      const int ks = (k < 0 && ext != 2) ? 0 : k, js = (j < 0 && ext != 1) ? 0 : j;
      double* dst = h + ((long)(k + 1) * n1 + (j + 1)) * n0;
      const double* src = a[ks][js];
      if (ext == 0) {
        memcpy(dst, src - 1, (size_t)(t[0] + 1) * sizeof(double));
      } else {
        dst[0] = src[0];
        memcpy(dst + 1, src, (size_t)t[0] * sizeof(double));
      }
    }
}

void unpack_ghosted(const double* h, double*** a, const int t[3], const int lo[3], const int hi[3])
{
  const long n0 = t[0] + 1, n1 = t[1] + 1;
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      memcpy(&a[k][j][lo[0]], h + ((long)(k + 1) * n1 + (j + 1)) * n0 + lo[0] + 1,
             (size_t)(hi[0] - lo[0] + 1) * sizeof(double));
}

// The face areas of a uniform Cartesian grid (set_geometry.c:146-210): one
// value per axis, checked over the whole array.
bool uniform_areas(const Grid* grid, BoxInfo& b, const char* fn)
{
  for (int a = 0; a < 3; a++) {
    if (!grid->A[a]) {
      report(fn, "grid->A is NULL");
      return false;
    }
    const double v = grid->A[a][0][0][0];
    for (int k = (a == 2 ? -1 : 0); k < b.t[2]; k++)
      for (int j = (a == 1 ? -1 : 0); j < b.t[1]; j++)
        for (int i = (a == 0 ? -1 : 0); i < b.t[0]; i++)
          if (grid->A[a][k][j][i] != v) {
            report(fn, "non-uniform face areas are not supported on the device");
            return false;
          }
    b.area[a] = v;
  }
  return true;
}

// Interior of Vc / Vs / Uc <-> packed host arrays [var][nz][ny][nx].
inline long pidx(const BoxInfo& b, int i, int j, int k)
{
  return ((long)k * b.n[1] + j) * b.n[0] + i;
}

// f(k0, k1) over the planes 0..nk-1, split over up to 16 host threads
// (one copy of a 256^3 array moves 134 MB; a single core copies ~20 GB/s,
// below the host link).
template <class F>
void for_planes(int nk, F f)
{
  static const int T = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int nt = std::min(T, std::max(1, nk / 4));
  if (nt <= 1) {
    f(0, nk);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt - 1);
  for (int t = 1; t < nt; t++) th.emplace_back(f, (int)((long)t * nk / nt), (int)((long)(t + 1) * nk / nt));
  f(0, nk / nt);
  for (auto& x : th) x.join();
}

// Variable q of Vc's interior into dst (n cells); returns whether a value
// was a NaN.
bool pack_prim_var(const BoxInfo& b, double**** Vc, int q, double* dst)
{
  std::vector<unsigned char> bad((size_t)b.n[2], 0);   // a NaN in plane k
  for_planes(b.n[2], [&](int k0, int k1) {
    for (int k = k0; k < k1; k++) {
      bool any = false;
      for (int j = 0; j < b.n[1]; j++) {
        const double* row = Vc[q][k + b.beg[2]][j + b.beg[1]] + b.beg[0];
        double* out = dst + pidx(b, 0, j, k);
        for (int i = 0; i < b.n[0]; i++) {
          const double x = row[i];
          out[i] = x;
          any |= (x != x);
        }
      }
      bad[k] = any;
    }
  });
  return std::any_of(bad.begin(), bad.end(), [](unsigned char c) { return c != 0; });
}

void unpack_prim_var(const BoxInfo& b, const double* src, double**** Vc, int q)
{
  for_planes(b.n[2], [&](int k0, int k1) {
    for (int k = k0; k < k1; k++)
      for (int j = 0; j < b.n[1]; j++) {
        double* row = Vc[q][k + b.beg[2]][j + b.beg[1]] + b.beg[0];
        memcpy(row, src + pidx(b, 0, j, k), sizeof(double) * b.n[0]);
      }
  });
}

void pack_prim(const BoxInfo& b, double**** Vc, std::vector<double>& h)
{
  const long n = (long)b.n[0] * b.n[1] * b.n[2];
  h.resize(8 * n);
  for (int q = 0; q < 8; q++) (void)pack_prim_var(b, Vc, q, h.data() + q * n);
}

void unpack_prim(const BoxInfo& b, const std::vector<double>& h, double**** Vc)
{
  const long n = (long)b.n[0] * b.n[1] * b.n[2];
  for (int q = 0; q < 8; q++) unpack_prim_var(b, h.data() + q * n, Vc, q);
}

// Vs[a][k][j][i] holds the face at +1/2 of cell (i,j,k) along axis a;
// axes beyond the box's dimensions pack as zeros.
void pack_face(const BoxInfo& b, double**** Vs, int a, double* dst)
{
  if (a >= b.dims) {
    memset(dst, 0, sizeof(double) * (size_t)b.n[0] * b.n[1] * b.n[2]);
    return;
  }
  for_planes(b.n[2], [&](int k0, int k1) {
    for (int k = k0; k < k1; k++)
      for (int j = 0; j < b.n[1]; j++) {
        const double* row = Vs[a][k + b.beg[2]][j + b.beg[1]] + b.beg[0];
        memcpy(dst + pidx(b, 0, j, k), row, sizeof(double) * b.n[0]);
      }
  });
}

// Interior faces of axis a, plus the low-side boundary faces that the
// reference's CT_Update also writes (ct_update.c:79-190), as their periodic
// images.
void unpack_face(const BoxInfo& b, const double* src, double**** Vs, int a)
{
  if (a >= b.dims) return;
  for_planes(b.n[2], [&](int k0, int k1) {
    for (int k = k0; k < k1; k++)
      for (int j = 0; j < b.n[1]; j++) {
        double* row = Vs[a][k + b.beg[2]][j + b.beg[1]] + b.beg[0];
        memcpy(row, src + pidx(b, 0, j, k), sizeof(double) * b.n[0]);
      }
  });
  if (a == 0)
    for (int k = 0; k < b.n[2]; k++)
      for (int j = 0; j < b.n[1]; j++)
        Vs[0][k + b.beg[2]][j + b.beg[1]][b.beg[0] - 1] = src[pidx(b, b.n[0] - 1, j, k)];
  if (a == 1)
    for (int k = 0; k < b.n[2]; k++)
      for (int i = 0; i < b.n[0]; i++)
        Vs[1][k + b.beg[2]][b.beg[1] - 1][i + b.beg[0]] = src[pidx(b, i, b.n[1] - 1, k)];
  if (a == 2)
    for (int j = 0; j < b.n[1]; j++)
      for (int i = 0; i < b.n[0]; i++)
        Vs[2][b.beg[2] - 1][j + b.beg[1]][i + b.beg[0]] = src[pidx(b, i, j, b.n[2] - 1)];
}

void pack_faces(const BoxInfo& b, double**** Vs, std::vector<double>& h)
{
  const long n = (long)b.n[0] * b.n[1] * b.n[2];
  h.resize(3 * n);
  for (int a = 0; a < 3; a++) pack_face(b, Vs, a, h.data() + a * n);
}

void unpack_faces(const BoxInfo& b, const std::vector<double>& h, double**** Vs)
{
  const long n = (long)b.n[0] * b.n[1] * b.n[2];
  for (int a = 0; a < 3; a++) unpack_face(b, h.data() + a * n, Vs, a);
}

// Uc[k][j][i][nv] <-> the device accumulator (rho, mx, my, mz, E).
const int kConsVar[5] = {RHO, VX1, VX2, VX3, ENG};

void pack_cons(const BoxInfo& b, double**** UU, std::vector<double>& h)
{
  const long n = (long)b.n[0] * b.n[1] * b.n[2];
  h.resize(5 * n);
  for (int k = 0; k < b.n[2]; k++)
    for (int j = 0; j < b.n[1]; j++)
      for (int i = 0; i < b.n[0]; i++) {
        const double* u = UU[k + b.beg[2]][j + b.beg[1]][i + b.beg[0]];
        for (int q = 0; q < 5; q++) h[q * n + pidx(b, i, j, k)] = u[kConsVar[q]];
      }
}

void unpack_cons(const BoxInfo& b, const std::vector<double>& h, double**** UU)
{
  const long n = (long)b.n[0] * b.n[1] * b.n[2];
  for (int k = 0; k < b.n[2]; k++)
    for (int j = 0; j < b.n[1]; j++)
      for (int i = 0; i < b.n[0]; i++) {
        double* u = UU[k + b.beg[2]][j + b.beg[1]][i + b.beg[0]];
        for (int q = 0; q < 5; q++) u[kConsVar[q]] = h[q * n + pidx(b, i, j, k)];
      }
}

bool upload(pgpu_domain* dom, double* const* dst, const std::vector<double>& h, int nvar)
{
  const size_t nb = (size_t)dom->g.n * sizeof(double);
  for (int q = 0; q < nvar; q++)
    if (hipMemcpyAsync(dst[q], h.data() + (size_t)q * dom->g.n, nb, hipMemcpyHostToDevice,
                       dom->stream) != hipSuccess)
      return false;
  return true;
}

bool download(pgpu_domain* dom, double* const* src, std::vector<double>& h, int nvar)
{
  const size_t nb = (size_t)dom->g.n * sizeof(double);
  h.resize((size_t)nvar * dom->g.n);
  for (int q = 0; q < nvar; q++)
    if (hipMemcpyAsync(h.data() + (size_t)q * dom->g.n, src[q], nb, hipMemcpyDeviceToHost,
                       dom->stream) != hipSuccess)
      return false;
  return hipStreamSynchronize(dom->stream) == hipSuccess;
}

bool read_maxima(pgpu_domain* dom, double& cdt_max, double& mach_max)
{
  unsigned long long v[2];
  if (hipMemcpyAsync(v, &dom->sc->invdt_max, sizeof(v), hipMemcpyDeviceToHost, dom->stream) !=
          hipSuccess ||
      hipStreamSynchronize(dom->stream) != hipSuccess)
    return false;
  memcpy(&cdt_max, &v[0], sizeof(double));
  memcpy(&mach_max, &v[1], sizeof(double));
  return true;
}


// ---------------------------------------------------------------------
// USERDEF sides (boundary.c:258-292): the program's UserDefBoundary on the
// host, on PLUTO's own arrays.
// ---------------------------------------------------------------------

constexpr int kSideX1Beg = 101;   // X1_BEG .. X3_END = 101 .. 106 (pluto.h:159-164)

// The centre box and the three face boxes of side `is` (boundary.c:144-165;
// a BEG side's boxes run backwards along the normal, as RBoxDefine gets them)
void side_boxes(const int t[3], const int beg[3], const int end[3], int is, RBox box[4])
{
  int lo[3] = {0, 0, 0}, hi[3] = {t[0] - 1, t[1] - 1, t[2] - 1};
  const int a = is / 2;
  if (is & 1) {
    lo[a] = end[a] + 1;
    hi[a] = t[a] - 1;
  } else {
    lo[a] = beg[a] - 1;
    hi[a] = 0;
  }
  for (int f = 0; f < 4; f++) {
    int l[3] = {lo[0], lo[1], lo[2]}, h[3] = {hi[0], hi[1], hi[2]};
    if (f > 0) {
      const int x = f - 1;
      if (l[x] < h[x]) l[x] -= 1;
      else h[x] -= 1;
    }
    RBox& r = box[f];
    memset(&r, 0, sizeof(r));
    r.ibeg = l[0];
    r.iend = h[0];
    r.jbeg = l[1];
    r.jend = h[1];
    r.kbeg = l[2];
    r.kend = h[2];
    r.vpos = f;   // CENTER, X1FACE, X2FACE, X3FACE
  }
}

// CT_AverageTransverseMagField (ct_field_average.c:275-457, Cartesian): the
// cell-centred field components tangent to the side, in its ghost cells
void transverse_average(const Data* d, const int t[3], const int beg[3], const int end[3], int is)
{
  const int a = is / 2;
  int lo[3] = {0, 0, 0}, hi[3] = {t[0] - 1, t[1] - 1, t[2] - 1};
  if (is & 1) lo[a] = end[a] + 1;
  else hi[a] = beg[a] - 1;
  double*** B[3] = {d->Vc[BX1], d->Vc[BX1 + 1], d->Vc[BX1 + 2]};
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++) {
        if (a != 0) B[0][k][j][i] = 0.5 * (d->Vs[0][k][j][i] + d->Vs[0][k][j][i - 1]);
        if (a != 1) B[1][k][j][i] = 0.5 * (d->Vs[1][k][j][i] + d->Vs[1][k][j - 1][i]);
        if (a != 2) B[2][k][j][i] = 0.5 * (d->Vs[2][k][j][i] + d->Vs[2][k - 1][j][i]);
      }
}

// The host part of a USERDEF side on PLUTO's arrays: the hook on the centre
// and face boxes (boundary.c:285-288), then the transverse average
// (boundary.c:290)
void userdef_on_host(const Data* d, Grid* grid, const int t[3], const int beg[3],
                     const int end[3], int is)
{
  RBox box[4];
  side_boxes(t, beg, end, is, box);
  for (int f = 0; f < 4; f++) UserDefBoundary(d, &box[f], kSideX1Beg + is, grid);
  transverse_average(d, t, beg, end, is);
}

// pgpu_userdef_fn for the ghosted domain behind AdvanceStep: the slab goes
// into PLUTO's own arrays, the hook runs there, and the slab comes back
struct UserdefCtx {
  Data* d;
  Grid* grid;
  int t[3], beg[3], end[3];
};

void slab_to_pluto(const pgpu_slab* s, const Data* d, bool to_pluto)
{
  for (int q = 0; q < 11; q++) {
    const int ext = q < 8 ? -1 : q - 8;   // the axis along which index -1 exists
    double* buf = q < 8 ? s->vc[q] : s->vs[q - 8];
    double*** arr = q < 8 ? d->Vc[q] : d->Vs[q - 8];
    for (int k = s->lo[2]; k <= s->hi[2]; k++) {
      if (k < 0 && ext != 2) continue;
      for (int j = s->lo[1]; j <= s->hi[1]; j++) {
        if (j < 0 && ext != 1) continue;
        const long row = ((long)(k - s->lo[2]) * s->ld[1] + (j - s->lo[1])) * s->ld[0];
        for (int i = s->lo[0]; i <= s->hi[0]; i++) {
          if (i < 0 && ext != 0) continue;
          double& h = buf[row + (i - s->lo[0])];
          if (to_pluto) arr[k][j][i] = h;
          else h = arr[k][j][i];
        }
      }
    }
  }
}

int userdef_adapter(void* ctx, int side, int stage, pgpu_slab* s)
{
  UserdefCtx* c = (UserdefCtx*)ctx;
  slab_to_pluto(s, c->d, true);
  g_intStage = stage;
  userdef_on_host(c->d, c->grid, c->t, c->beg, c->end, side);
  slab_to_pluto(s, c->d, false);
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------
// Row-level entry points
// ---------------------------------------------------------------------

// One Riemann solver over the interfaces beg..end of a sweep row (the
// body of HLLD_Solver and HLL_Solver).
static void riemann_row(int solver, const char* who, const Sweep* sweep, int beg, int end,
                        double* cmax)
{
  const int n = end - beg + 1;
  if (n <= 0) return;
  const int dir = g_dir;
  Scratch& S = scratch();
  double* d = S.get((size_t)n * (3 * NV + 5) * sizeof(double));
  if (!d) {
    report(who, "device scratch allocation failed");
    return;
  }
  double *dvl = d, *dvr = d + (size_t)n * NV, *dfl = d + 2 * (size_t)n * NV;
  double *dp = d + 3 * (size_t)n * NV, *dsl = dp + n, *dsr = dsl + n, *dcm = dsr + n, *dmach = dcm + n;
  std::vector<double> h((size_t)n * (3 * NV + 5));
  for (int i = 0; i < n; i++)
    for (int q = 0; q < NV; q++) {
      h[(size_t)i * NV + q] = sweep->stateL.v[beg + i][q];
      h[(size_t)(n + i) * NV + q] = sweep->stateR.v[beg + i][q];
    }
  bool ok = hipMemcpyAsync(d, h.data(), 2 * (size_t)n * NV * sizeof(double), hipMemcpyHostToDevice,
                           S.s) == hipSuccess;
  ok = ok && launch_riemann_rows(solver, dir, n, dvl, dvr, g_gamma, dfl, dp, dsl, dsr, dcm, dmach,
                                 S.s) == 0;
  ok = ok && hipMemcpyAsync(h.data() + 2 * (size_t)n * NV, dfl,
                            ((size_t)n * NV + 5 * (size_t)n) * sizeof(double),
                            hipMemcpyDeviceToHost, S.s) == hipSuccess;
  ok = ok && hipStreamSynchronize(S.s) == hipSuccess;
  if (!ok) {
    report(who, "device launch failed");
    return;
  }
  const double* f = h.data() + 2 * (size_t)n * NV;
  const double* r = f + (size_t)n * NV;
  for (int i = 0; i < n; i++) {
    for (int q = 0; q < NV; q++) sweep->flux[beg + i][q] = f[(size_t)i * NV + q];
    sweep->press[beg + i] = r[i];
    sweep->SL[beg + i] = r[n + i];
    sweep->SR[beg + i] = r[2 * n + i];
    cmax[beg + i] = r[3 * n + i];
    const double m = r[4 * n + i];
    g_maxMach = (m >= g_maxMach) ? m : g_maxMach;   // MAX(scrh, g_maxMach)
  }
}

extern "C" void HLLD_Solver(const Sweep* sweep, int beg, int end, double* cmax, Grid* grid)
{
  (void)grid;
  riemann_row(SOLVER_HLLD, "HLLD_Solver", sweep, beg, end, cmax);
}

extern "C" void HLL_Solver(const Sweep* sweep, int beg, int end, double* cmax, Grid* grid)
{
  (void)grid;
  riemann_row(SOLVER_HLL, "HLL_Solver", sweep, beg, end, cmax);
}

extern "C" void HLLC_Solver(const Sweep* sweep, int beg, int end, double* cmax, Grid* grid)
{
  (void)grid;
  riemann_row(SOLVER_HLLC, "HLLC_Solver", sweep, beg, end, cmax);
}

extern "C" void LF_Solver(const Sweep* sweep, int beg, int end, double* cmax, Grid* grid)
{
  (void)grid;
  riemann_row(SOLVER_TVDLF, "LF_Solver", sweep, beg, end, cmax);
}

extern "C" void Roe_Solver(const Sweep* sweep, int beg, int end, double* cmax, Grid* grid)
{
  (void)grid;
  riemann_row(SOLVER_ROE, "Roe_Solver", sweep, beg, end, cmax);
}

extern "C" void States(const Sweep* sweep, int beg, int end, Grid* grid)
{
  (void)grid;
  const int n = end - beg + 1;
  if (n <= 0) return;
  const int dir = g_dir;
  const int BN = BX1 + dir;
  Scratch& S = scratch();
  const size_t nin = (size_t)(n + 2) * NV + (n + 1);
  double* d = S.get((nin + 4 * (size_t)n * NV) * sizeof(double));
  if (!d) {
    report("States", "device scratch allocation failed");
    return;
  }
  std::vector<double> h(nin + 4 * (size_t)n * NV);
  for (int i = 0; i < n + 2; i++)
    for (int q = 0; q < NV; q++) h[(size_t)i * NV + q] = sweep->stateC.v[beg - 1 + i][q];
  for (int i = 0; i < n + 1; i++) h[(size_t)(n + 2) * NV + i] = sweep->bn[beg - 1 + i];
  double* dv = d;
  double* dbn = d + (size_t)(n + 2) * NV;
  double* out = d + nin;
  bool ok = hipMemcpyAsync(d, h.data(), nin * sizeof(double), hipMemcpyHostToDevice, S.s) ==
            hipSuccess;
  ok = ok && launch_plm_rows(dir, n, dv, dbn, g_gamma, out, out + (size_t)n * NV,
                             out + 2 * (size_t)n * NV, out + 3 * (size_t)n * NV, S.s) == 0;
  ok = ok && hipMemcpyAsync(h.data() + nin, out, 4 * (size_t)n * NV * sizeof(double),
                            hipMemcpyDeviceToHost, S.s) == hipSuccess;
  ok = ok && hipStreamSynchronize(S.s) == hipSuccess;
  if (!ok) {
    report("States", "device launch failed");
    return;
  }
  const double* vp = h.data() + nin;
  const double* vm = vp + (size_t)n * NV;
  const double* up = vm + (size_t)n * NV;
  const double* um = up + (size_t)n * NV;
  for (int i = 0; i < n; i++)
    for (int q = 0; q < NV; q++) {
      sweep->stateL.v[beg + i][q] = vp[(size_t)i * NV + q];
      sweep->stateR.v[beg + i - 1][q] = vm[(size_t)i * NV + q];
      sweep->stateL.u[beg + i][q] = up[(size_t)i * NV + q];
      sweep->stateR.u[beg + i - 1][q] = um[(size_t)i * NV + q];
    }
  // the face loop of plm_states.c:272-274 also touches the two end states
  sweep->stateL.v[beg - 1][BN] = sweep->bn[beg - 1];
  sweep->stateR.v[end][BN] = sweep->bn[end];
}

extern "C" void PrimToCons(double** uprim, double** ucons, int ibeg, int iend)
{
  const int n = iend - ibeg + 1;
  if (n <= 0) return;
  Scratch& S = scratch();
  double* d = S.get(2 * (size_t)n * NV * sizeof(double));
  if (!d) {
    report("PrimToCons", "device scratch allocation failed");
    return;
  }
  std::vector<double> h(2 * (size_t)n * NV);
  for (int i = 0; i < n; i++)
    for (int q = 0; q < NV; q++) h[(size_t)i * NV + q] = uprim[ibeg + i][q];
  bool ok = hipMemcpyAsync(d, h.data(), (size_t)n * NV * sizeof(double), hipMemcpyHostToDevice,
                           S.s) == hipSuccess;
  ok = ok && launch_p2c_rows(n, d, g_gamma, d + (size_t)n * NV, S.s) == 0;
  ok = ok && hipMemcpyAsync(h.data() + (size_t)n * NV, d + (size_t)n * NV,
                            (size_t)n * NV * sizeof(double), hipMemcpyDeviceToHost, S.s) ==
                 hipSuccess;
  ok = ok && hipStreamSynchronize(S.s) == hipSuccess;
  if (!ok) {
    report("PrimToCons", "device launch failed");
    return;
  }
  for (int i = 0; i < n; i++)
    for (int q = 0; q < NV; q++) ucons[ibeg + i][q] = h[(size_t)(n + i) * NV + q];
}

extern "C" int ConsToPrim(double** ucons, double** uprim, int ibeg, int iend,
                          unsigned char* flag)
{
  const int n = iend - ibeg + 1;
  if (n <= 0) return 0;
  Scratch& S = scratch();
  double* d = S.get((2 * (size_t)n * NV + (n + 7) / 8 + 1) * sizeof(double));
  if (!d) {
    report("ConsToPrim", "device scratch allocation failed");
    return 1;
  }
  std::vector<double> h(2 * (size_t)n * NV);
  std::vector<unsigned char> f(n);
  for (int i = 0; i < n; i++)
    for (int q = 0; q < NV; q++) h[(size_t)i * NV + q] = ucons[ibeg + i][q];
  unsigned char* dflag = (unsigned char*)(d + 2 * (size_t)n * NV);
  bool ok = hipMemcpyAsync(d, h.data(), (size_t)n * NV * sizeof(double), hipMemcpyHostToDevice,
                           S.s) == hipSuccess;
  ok = ok && launch_c2p_rows(n, d, g_gamma, d + (size_t)n * NV, dflag, S.s) == 0;
  ok = ok && hipMemcpyAsync(h.data(), d, 2 * (size_t)n * NV * sizeof(double),
                            hipMemcpyDeviceToHost, S.s) == hipSuccess;
  ok = ok && hipMemcpyAsync(f.data(), dflag, n, hipMemcpyDeviceToHost, S.s) == hipSuccess;
  ok = ok && hipStreamSynchronize(S.s) == hipSuccess;
  if (!ok) {
    report("ConsToPrim", "device launch failed");
    return 1;
  }
  int ifail = 0;
  for (int i = 0; i < n; i++) {
    for (int q = 0; q < NV; q++) {
      ucons[ibeg + i][q] = h[(size_t)i * NV + q];
      uprim[ibeg + i][q] = h[(size_t)(n + i) * NV + q];
    }
    if (f[i]) {
      flag[ibeg + i] |= f[i];   // absolute index, as mappers.c:120-134
      ifail = 1;
    }
  }
  return ifail;
}

extern "C" void CT_AverageMagneticField(double**** bf, double**** UU, Grid* grid)
{
  BoxInfo b;
  b.dims = (grid->np_int[2] > 1) ? 3 : 2;
  for (int a = 0; a < 3; a++) {
    b.beg[a] = grid->lbeg[a];
    b.end[a] = grid->lend[a];
  }
  // domain plus one ghost layer in each active direction
  const int off[3] = {1, 1, b.dims == 3 ? 1 : 0};
  int lo[3], m[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = (a == 2 && b.dims == 2) ? b.beg[2] : b.beg[a] - off[a];
    m[a] = (a == 2 && b.dims == 2) ? 1 : (b.end[a] - b.beg[a] + 1) + 2 * off[a];
  }
  const size_t nfx = (size_t)(m[0] + 1) * m[1] * m[2];
  const size_t nfy = (size_t)m[0] * (m[1] + 1) * m[2];
  const size_t nfz = (size_t)m[0] * m[1] * (m[2] + 1);
  const size_t nout = 3 * (size_t)m[0] * m[1] * m[2];
  std::vector<double> h(nfx + nfy + nfz + nout, 0.0);
  double *hx = h.data(), *hy = hx + nfx, *hz = hy + nfy;
  for (int k = 0; k < m[2]; k++)
    for (int j = 0; j < m[1]; j++)
      for (int i = 0; i <= m[0]; i++)
        hx[((size_t)k * m[1] + j) * (m[0] + 1) + i] = bf[0][lo[2] + k][lo[1] + j][lo[0] + i - 1];
  for (int k = 0; k < m[2]; k++)
    for (int j = 0; j <= m[1]; j++)
      for (int i = 0; i < m[0]; i++)
        hy[((size_t)k * (m[1] + 1) + j) * m[0] + i] = bf[1][lo[2] + k][lo[1] + j - 1][lo[0] + i];
  if (b.dims == 3)
    for (int k = 0; k <= m[2]; k++)
      for (int j = 0; j < m[1]; j++)
        for (int i = 0; i < m[0]; i++)
          hz[((size_t)k * m[1] + j) * m[0] + i] = bf[2][lo[2] + k - 1][lo[1] + j][lo[0] + i];
  Scratch& S = scratch();
  double* d = S.get(h.size() * sizeof(double));
  if (!d) {
    report("CT_AverageMagneticField", "device scratch allocation failed");
    return;
  }
  bool ok = hipMemcpyAsync(d, h.data(), (nfx + nfy + nfz) * sizeof(double),
                           hipMemcpyHostToDevice, S.s) == hipSuccess;
  ok = ok && launch_face_average_box(m[0], m[1], m[2], d, d + nfx, d + nfx + nfy, b.dims,
                                     d + nfx + nfy + nfz, S.s) == 0;
  ok = ok && hipMemcpyAsync(h.data() + nfx + nfy + nfz, d + nfx + nfy + nfz,
                            nout * sizeof(double), hipMemcpyDeviceToHost, S.s) == hipSuccess;
  ok = ok && hipStreamSynchronize(S.s) == hipSuccess;
  if (!ok) {
    report("CT_AverageMagneticField", "device launch failed");
    return;
  }
  const double* o = h.data() + nfx + nfy + nfz;
  for (int k = 0; k < m[2]; k++)
    for (int j = 0; j < m[1]; j++)
      for (int i = 0; i < m[0]; i++) {
        const size_t c = ((size_t)k * m[1] + j) * m[0] + i;
        double* u = UU[lo[2] + k][lo[1] + j][lo[0] + i];
        u[BX1] = o[3 * c + 0];
        u[BX2] = o[3 * c + 1];
        if (b.dims == 3) u[BX3] = o[3 * c + 2];
      }
}

extern "C" Riemann_Solver* SetSolver(const char* solver)
{
  if (solver && strcmp(solver, "hlld") == 0) return &HLLD_Solver;
  if (solver && strcmp(solver, "hll") == 0) return &HLL_Solver;
  if (solver && strcmp(solver, "hllc") == 0) return &HLLC_Solver;
  if (solver && strcmp(solver, "tvdlf") == 0) return &LF_Solver;
  if (solver && strcmp(solver, "roe") == 0) return &Roe_Solver;
  report("SetSolver", "only the \"hlld\", \"hll\", \"hllc\", \"tvdlf\" and \"roe\" solvers are provided by this library");
  return nullptr;
}

extern "C" void pgpu_finalize(void)
{
  if (g_box_dom) pgpu_domain_destroy(g_box_dom);
  g_box_dom = nullptr;
  release_scratch();
}

// EMF average of AdvanceStep (PLUTO's CT_EMF_AVERAGE, a compile-time macro
// in the reference, ct.h:43-45)
static int g_emf_average = PGPU_EMF_UCT_HLL;

extern "C" int PGPU_SetEMFAverage(int ct_emf_average)
{
  if (ct_emf_average < PGPU_EMF_ARITHMETIC || ct_emf_average > PGPU_EMF_UCT_HLL) {
    report("PGPU_SetEMFAverage", "unknown CT_EMF_AVERAGE label");
    return 1;
  }
  g_emf_average = ct_emf_average;
  return 0;
}

// CheckNaN (tools.c:20-38) stops a serial PLUTO run with exit status 1
// (QUIT_PLUTO, macros.h:169) when a stage input holds a NaN.  The library
// does not exit: the whole-box entry points report it, as an error that
// changes nothing (AdvanceStep returns nonzero), for the input they receive
// and, in AdvanceStep, for the input of the second stage.
static void report_nan(const char* fn)
{
  report(fn, "NaN found in a stage input (CheckNaN; PLUTO stops the run here)");
}

static bool has_nan_host(const std::vector<double>& h)
{
  for (double x : h)
    if (x != x) return true;
  return false;
}

static bool device_nans(pgpu_domain* dom)
{
  unsigned int nans = 0;
  return hipMemcpy(&nans, &dom->sc->nan_cells, sizeof(nans), hipMemcpyDeviceToHost) ==
             hipSuccess &&
         nans != 0;
}

// PGPU_SOLVER_* of a Riemann_Solver pointer, -1 when it is not one of ours
static int solver_of(Riemann_Solver* r)
{
  if (r == &HLLD_Solver) return PGPU_SOLVER_HLLD;
  if (r == &HLL_Solver) return PGPU_SOLVER_HLL;
  if (r == &HLLC_Solver) return PGPU_SOLVER_HLLC;
  if (r == &LF_Solver) return PGPU_SOLVER_TVDLF;
  if (r == &Roe_Solver) return PGPU_SOLVER_ROE;
  return -1;
}

// ---------------------------------------------------------------------
// Whole-box entry points
// ---------------------------------------------------------------------

// AdvanceStep on a box with non-periodic sides: the whole ghosted box goes
// to the device, every stage starts with Boundary()'s ghost fill there
// (rk_step.c:81,151), and the results come back over the ranges the
// reference's step writes: Vc and flag over the domain (ConsToPrim3D,
// rk_step.c:136,189), Vs over the CT_Update ranges (ct_update.c:79-190).
// The ghost zones keep their host values; the next step fills them again.
static int advance_ghosted(Data* d, timeStep* Dts, const BoxInfo& b, pgpu_domain* dom,
                           Grid* grid)
{
  const char* fn = "AdvanceStep";
  const long n = dom->g.n;
  std::vector<double> hp(8 * n), hf(3 * n);
  for (int q = 0; q < 8; q++) pack_ghosted(hp.data() + q * n, d->Vc[q], b.t, -1);
  for (int a = 0; a < 3; a++) pack_ghosted(hf.data() + a * n, d->Vs[a], b.t, a);
  for (int q = 0; q < 8; q++)
    for (int k = b.beg[2]; k <= b.end[2]; k++)
      for (int j = b.beg[1]; j <= b.end[1]; j++)
        for (int i = b.beg[0]; i <= b.end[0]; i++)
          if (isnan(d->Vc[q][k][j][i])) {
            report_nan(fn);
            return 1;
          }
  if (hipMemsetAsync(&dom->sc->nan_cells, 0, 2 * sizeof(unsigned int), dom->stream) != hipSuccess ||
      !upload(dom, dom->V, hp, 8) || !upload(dom, dom->bA, hf, 3)) {
    report(fn, "upload failed");
    return 1;
  }
  // USERDEF sides: the program's UserDefBoundary runs on PLUTO's arrays,
  // on the slab of its side, before each stage (inside pgpu_run)
  UserdefCtx uctx;
  uctx.d = d;
  uctx.grid = grid;
  for (int a = 0; a < 3; a++) {
    uctx.t[a] = b.t[a];
    uctx.beg[a] = b.beg[a];
    uctx.end[a] = b.end[a];
  }
  bool has_userdef = false;
  for (int is = 0; is < 6; is++) has_userdef = has_userdef || b.type[is] == BOUND_USERDEF;
  pgpu_set_userdef_boundary(dom, has_userdef ? userdef_adapter : nullptr, &uctx);
  const bool ran = pgpu_set_dt(dom, g_dt, 1.0, 1.0, 1) == PGPU_OK && pgpu_run(dom, 1) == PGPU_OK;
  pgpu_set_userdef_boundary(dom, nullptr, nullptr);
  if (!ran) {
    report(fn, pgpu_last_error());
    return 1;
  }
  double sc[6];
  double cdt_raw = 0.0;
  std::vector<unsigned char> fl(n);
  if (!download(dom, dom->V, hp, 8) || !download(dom, dom->bA, hf, 3) ||
      pgpu_get_scalars(dom, sc) != PGPU_OK ||
      hipMemcpy(&cdt_raw, &dom->sc->cdt_raw_last, sizeof(double), hipMemcpyDeviceToHost) !=
          hipSuccess ||
      hipMemcpy(fl.data(), dom->flag, (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) {
    report(fn, "download failed");
    return 1;
  }
  if (device_nans(dom)) {
    report_nan(fn);
    return 1;
  }
  for (int q = 0; q < 8; q++) unpack_ghosted(hp.data() + q * n, d->Vc[q], b.t, b.beg, b.end);
  for (int a = 0; a < 3; a++) {
    int lo[3] = {b.beg[0], b.beg[1], b.beg[2]};
    lo[a] -= 1;
    unpack_ghosted(hf.data() + a * n, d->Vs[a], b.t, lo, b.end);
  }
  if (d->flag) {
    const long n0 = b.t[0] + 1, n1 = b.t[1] + 1;
    for (int k = b.beg[2]; k <= b.end[2]; k++)
      for (int j = b.beg[1]; j <= b.end[1]; j++)
        for (int i = b.beg[0]; i <= b.end[0]; i++)
          d->flag[k][j][i] |= fl[((long)(k + 1) * n1 + (j + 1)) * n0 + i + 1];
  }
  // update_stage.c:310-311 and hll_speed.c:105
  double inv = Dts->invDt_hyp;
  inv = (inv >= cdt_raw) ? inv : cdt_raw;
  Dts->invDt_hyp = inv / (double)b.dims;
  g_maxMach = (sc[2] >= g_maxMach) ? sc[2] : g_maxMach;
  g_intStage = 2;
  return 0;
}

extern "C" int AdvanceStep(Data* d, Riemann_Solver* Riemann, timeStep* Dts, Grid* grid)
{
  const int solver = solver_of(Riemann);
  if (solver < 0) {
    report("AdvanceStep", "the device step implements the HLLD, HLL, HLLC, LF and Roe solvers only");
    return 1;
  }
  BoxInfo b;
  if (!box_of(grid, b, "AdvanceStep", true)) return 1;
  if (b.ghosted && !uniform_areas(grid, b, "AdvanceStep")) return 1;
  pgpu_domain* dom = domain_for(b, "AdvanceStep");
  if (!dom || pgpu_set_solver(dom, solver) != PGPU_OK) return 1;
  if (pgpu_set_emf_average(dom, g_emf_average) != PGPU_OK) {
    report("AdvanceStep", pgpu_last_error());
    return 1;
  }
  if (b.ghosted) return advance_ghosted(d, Dts, b, dom, grid);
  using clock = std::chrono::steady_clock;
  const auto c0 = clock::now();
  Staging& S = staging();
  const size_t n = (size_t)dom->g.n, nb = n * sizeof(double);
  if (!S.get(n)) {
    report("AdvanceStep", "pinned host staging allocation failed");
    return 1;
  }
  hipStream_t st = dom->stream;
  // pack each array into pinned memory and enqueue its upload before
  // packing the next; the step takes Vc as given (its cell-centred field
  // included), as PrimToCons3D does at rk_step.c:93
  double pack_s = 0.0;
  bool nan = false, ok = hipEventRecord(S.t[0], st) == hipSuccess;
  for (int q = 0; q < 11 && ok; q++) {
    const auto p0 = clock::now();
    double* h = S.h + (size_t)q * n;
    if (q < 8) nan |= pack_prim_var(b, d->Vc, q, h);
    else pack_face(b, d->Vs, q - 8, h);
    pack_s += std::chrono::duration<double>(clock::now() - p0).count();
    ok = hipMemcpyAsync(q < 8 ? dom->V[q] : dom->bA[q - 8], h, nb, hipMemcpyHostToDevice, st) ==
         hipSuccess;
  }
  ok = ok && hipEventRecord(S.t[1], st) == hipSuccess;
  if (!ok || nan) {
    (void)hipStreamSynchronize(st);
    if (nan) report_nan("AdvanceStep");
    else report("AdvanceStep", "upload failed");
    return 1;
  }
  if (pgpu_set_dt(dom, g_dt, 1.0, 1.0, 1) != PGPU_OK || pgpu_run(dom, 1) != PGPU_OK) {
    report("AdvanceStep", pgpu_last_error());
    return 1;
  }
  // the scalars, then the arrays and the flags, one event each
  ok = hipEventRecord(S.t[2], st) == hipSuccess &&
       hipMemcpyAsync(S.sc, dom->sc, offsetof(DevScalars, dt_hist), hipMemcpyDeviceToHost, st) ==
           hipSuccess &&
       hipEventRecord(S.ev[12], st) == hipSuccess;
  for (int q = 0; q < 11 && ok; q++)
    ok = hipMemcpyAsync(S.h + (size_t)q * n, q < 8 ? dom->V[q] : dom->bA[q - 8], nb,
                        hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipEventRecord(S.ev[q], st) == hipSuccess;
  ok = ok && hipMemcpyAsync(S.fl, dom->flag, n, hipMemcpyDeviceToHost, st) == hipSuccess &&
       hipEventRecord(S.ev[11], st) == hipSuccess && hipEventRecord(S.t[3], st) == hipSuccess;
  if (!ok || hipEventSynchronize(S.ev[12]) != hipSuccess) {
    (void)hipStreamSynchronize(st);
    report("AdvanceStep", "download failed");
    return 1;
  }
  if (S.sc->nan_cells) {
    (void)hipStreamSynchronize(st);
    report_nan("AdvanceStep");
    return 1;
  }
  // unpack each array as soon as its copy has landed, while the next copies
  double unpack_s = 0.0;
  for (int q = 0; q < 12 && ok; q++) {
    ok = hipEventSynchronize(S.ev[q]) == hipSuccess;
    if (!ok) break;
    const auto u0 = clock::now();
    if (q < 8) unpack_prim_var(b, S.h + (size_t)q * n, d->Vc, q);
    else if (q < 11) unpack_face(b, S.h + (size_t)q * n, d->Vs, q - 8);
    else if (d->flag)
      for_planes(b.n[2], [&](int k0, int k1) {
        for (int k = k0; k < k1; k++)
          for (int j = 0; j < b.n[1]; j++) {
            unsigned char* row = d->flag[k + b.beg[2]][j + b.beg[1]] + b.beg[0];
            const unsigned char* f = S.fl + pidx(b, 0, j, k);
            for (int i = 0; i < b.n[0]; i++) row[i] |= f[i];
          }
      });
    unpack_s += std::chrono::duration<double>(clock::now() - u0).count();
  }
  if (!ok || hipEventSynchronize(S.t[3]) != hipSuccess) {
    report("AdvanceStep", "download failed");
    return 1;
  }
  // update_stage.c:310-311 and hll_speed.c:105
  const double cdt_raw = S.sc->cdt_raw_last;
  double inv = Dts->invDt_hyp;
  inv = (inv >= cdt_raw) ? inv : cdt_raw;
  Dts->invDt_hyp = inv / (double)b.dims;
  g_maxMach = (S.sc->mach_last >= g_maxMach) ? S.sc->mach_last : g_maxMach;
  g_intStage = 2;
  float ms[3] = {0.f, 0.f, 0.f};
  for (int i = 0; i < 3; i++) (void)hipEventElapsedTime(&ms[i], S.t[i], S.t[i + 1]);
  g_times.pack += pack_s;
  g_times.h2d += 1e-3 * ms[0];
  g_times.step += 1e-3 * ms[1];
  g_times.d2h += 1e-3 * ms[2];
  g_times.unpack += unpack_s;
  g_times.total += std::chrono::duration<double>(clock::now() - c0).count();
  g_times.calls += 1;
  return 0;
}

extern "C" int PGPU_DropinTimes(pgpu_dropin_times* out, int reset)
{
  if (out) *out = g_times;
  if (reset) g_times = pgpu_dropin_times{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
  return 0;
}

// UpdateStage on a box with non-periodic sides, whose ghost zones the
// caller's Boundary() has filled (rk_step.c:81,151): the whole ghosted Vc,
// Vs and the domain's UU go to the device; UU comes back over the domain
// (update_stage.c:214-216) and Vs over the CT_Update ranges.
static void update_stage_ghosted(Data* d, Data_Arr UU, double dt, timeStep* Dts, const BoxInfo& b,
                                 pgpu_domain* dom)
{
  const char* fn = "UpdateStage";
  const long n = dom->g.n;
  const long n0 = b.t[0] + 1, n1 = b.t[1] + 1;
  auto at = [&](int i, int j, int k) { return ((long)(k + 1) * n1 + (j + 1)) * n0 + i + 1; };
  std::vector<double> hp(8 * n), hf(3 * n), hu(5 * n, 0.0);
  for (int q = 0; q < 8; q++) pack_ghosted(hp.data() + q * n, d->Vc[q], b.t, -1);
  for (int a = 0; a < 3; a++) pack_ghosted(hf.data() + a * n, d->Vs[a], b.t, a);
  for (int q = 0; q < 8; q++)
    for (int k = 0; k < b.t[2]; k++)
      for (int j = 0; j < b.t[1]; j++)
        for (int i = 0; i < b.t[0]; i++)
          if (isnan(hp[q * n + at(i, j, k)])) {
            report_nan(fn);
            return;
          }
  for (int k = b.beg[2]; k <= b.end[2]; k++)
    for (int j = b.beg[1]; j <= b.end[1]; j++)
      for (int i = b.beg[0]; i <= b.end[0]; i++)
        for (int q = 0; q < 5; q++) hu[q * n + at(i, j, k)] = UU[k][j][i][kConsVar[q]];
  if (!upload(dom, dom->V, hp, 8) || !upload(dom, dom->bA, hf, 3) || !upload(dom, dom->U, hu, 5) ||
      pgpu_set_dt(dom, dt, 1.0, 1.0, 1) != PGPU_OK) {
    report(fn, "upload failed");
    return;
  }
  StageIn in;
  for (int q = 0; q < 8; q++) in.V[q] = dom->V[q];
  for (int a = 0; a < 3; a++) in.b[a] = dom->bA[a];
  Cons U;
  for (int q = 0; q < 5; q++) U.u[q] = dom->U[q];
  FaceIn fA;
  FaceOut oA;
  for (int a = 0; a < 3; a++) {
    fA.b[a] = dom->bA[a];
    oA.b[a] = dom->bA[a];
  }
  const bool stage1 = (g_intStage == 1);
  int rc = launch_sweeps(dom->g, in, dom->sw, dom->sc, dom->stream, nullptr);
  if (!rc) rc = launch_update(dom->g, false, stage1, in, dom->sw, U, dom->sc, dom->stream, nullptr);
  if (!rc) rc = launch_emf_ct(dom->g, false, in, dom->sw, dom->e, fA, fA, oA, dom->sc, dom->stream,
                              nullptr);
  double cdt_max = 0.0, mach = 0.0;
  if (rc || !download(dom, dom->U, hu, 5) || !download(dom, dom->bA, hf, 3) ||
      !read_maxima(dom, cdt_max, mach)) {
    report(fn, "device stage failed");
    return;
  }
  for (int k = b.beg[2]; k <= b.end[2]; k++)
    for (int j = b.beg[1]; j <= b.end[1]; j++)
      for (int i = b.beg[0]; i <= b.end[0]; i++)
        for (int q = 0; q < 5; q++) UU[k][j][i][kConsVar[q]] = hu[q * n + at(i, j, k)];
  for (int a = 0; a < 3; a++) {
    int lo[3] = {b.beg[0], b.beg[1], b.beg[2]};
    lo[a] -= 1;
    unpack_ghosted(hf.data() + a * n, d->Vs[a], b.t, lo, b.end);
  }
  if (stage1) {
    double inv = Dts->invDt_hyp;
    inv = (inv >= cdt_max) ? inv : cdt_max;
    Dts->invDt_hyp = inv / (double)b.dims;
  }
  g_maxMach = (mach >= g_maxMach) ? mach : g_maxMach;
}

extern "C" void UpdateStage(Data* d, Data_Arr UU, double** aflux, Riemann_Solver* Riemann,
                            double dt, timeStep* Dts, Grid* grid)
{
  if (aflux != nullptr) {
    report("UpdateStage", "AMR flux storage (aflux) is not supported");
    return;
  }
  const int solver = solver_of(Riemann);
  if (solver < 0) {
    report("UpdateStage", "the device stage implements the HLLD, HLL, HLLC, LF and Roe solvers only");
    return;
  }
  if (g_emf_average != PGPU_EMF_UCT_HLL) {
    report("UpdateStage", "the single-stage entry point uses the UCT_HLL EMF average; "
                          "use AdvanceStep for the others");
    return;
  }
  BoxInfo b;
  if (!box_of(grid, b, "UpdateStage", true)) return;
  if (b.ghosted)
    for (int a = 0; a < 3; a++) b.area[a] = 1.0;   // no ghost fill here: the caller's Boundary()
  pgpu_domain* dom = domain_for(b, "UpdateStage");
  if (!dom || pgpu_set_solver(dom, solver) != PGPU_OK) return;
  if (b.ghosted) {
    update_stage_ghosted(d, UU, dt, Dts, b, dom);
    return;
  }
  std::vector<double> hp, hf, hu;
  pack_prim(b, d->Vc, hp);
  if (has_nan_host(hp)) {
    report_nan("UpdateStage");
    return;
  }
  pack_faces(b, d->Vs, hf);
  pack_cons(b, UU, hu);
  if (!upload(dom, dom->V, hp, 8) || !upload(dom, dom->bA, hf, 3) || !upload(dom, dom->U, hu, 5) ||
      pgpu_set_dt(dom, dt, 1.0, 1.0, 1) != PGPU_OK) {
    report("UpdateStage", "upload failed");
    return;
  }
  StageIn in;
  for (int q = 0; q < 8; q++) in.V[q] = dom->V[q];
  for (int a = 0; a < 3; a++) in.b[a] = dom->bA[a];
  Cons U;
  for (int q = 0; q < 5; q++) U.u[q] = dom->U[q];
  FaceIn fA;
  FaceOut oA;
  for (int a = 0; a < 3; a++) {
    fA.b[a] = dom->bA[a];
    oA.b[a] = dom->bA[a];
  }
  const bool stage1 = (g_intStage == 1);
  int rc = launch_sweeps(dom->g, in, dom->sw, dom->sc, dom->stream, nullptr);
  if (!rc) rc = launch_update(dom->g, false, stage1, in, dom->sw, U, dom->sc, dom->stream, nullptr);
  if (!rc) rc = launch_emf_ct(dom->g, false, in, dom->sw, dom->e, fA, fA, oA, dom->sc, dom->stream,
                              nullptr);
  double cdt_max = 0.0, mach = 0.0;
  if (rc || !download(dom, dom->U, hu, 5) || !download(dom, dom->bA, hf, 3) ||
      !read_maxima(dom, cdt_max, mach)) {
    report("UpdateStage", "device stage failed");
    return;
  }
  unpack_cons(b, hu, UU);
  unpack_faces(b, hf, d->Vs);
  if (stage1) {
    double inv = Dts->invDt_hyp;
    inv = (inv >= cdt_max) ? inv : cdt_max;
    Dts->invDt_hyp = inv / (double)b.dims;
  }
  g_maxMach = (mach >= g_maxMach) ? mach : g_maxMach;
}

// ---------------------------------------------------------------------
// CT entry points on PLUTO's arrays (MHD/CT/ct.h:94-112), UCT_HLL.  The
// row-level stores copy a sweep row into the EMF arrays on the host; the
// box-level EMF and face update run on the device over the EMF ranges the
// stores recorded.
// ---------------------------------------------------------------------

namespace {

double*** alloc3(long nz, long ny, long nx)
{
  double* block = (double*)calloc((size_t)(nz * ny * nx), sizeof(double));
  double** rows = (double**)malloc((size_t)(nz * ny) * sizeof(double*));
  double*** planes = (double***)malloc((size_t)nz * sizeof(double**));
  if (!block || !rows || !planes) return nullptr;
  for (long k = 0; k < nz; k++) {
    planes[k] = rows + k * ny;
    for (long j = 0; j < ny; j++) rows[k * ny + j] = block + (k * ny + j) * nx;
  }
  return planes;
}

// box [n2][n1][n0] <-> PLUTO 3D array a[k][j][i] (i from 0)
void pack3(double* h, double*** a, const int n[3])
{
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++) memcpy(h + ((size_t)k * n[1] + j) * n[0], a[k][j], n[0] * sizeof(double));
}

void unpack3_range(const double* h, double*** a, const int n[3], const int lo[3], const int hi[3])
{
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++) a[k][j][i] = h[((size_t)k * n[1] + j) * n[0] + i];
}

bool ct_box(const Grid* grid, const EMF* emf, const char* fn, int n[3], int lo[3], int hi[3])
{
  if (grid->np_int[2] <= 1) {
    report(fn, "the PLUTO entry points support DIMENSIONS 3 (COMPONENTS 3) builds");
    return false;
  }
  if (!emf) {
    report(fn, "data->emf is NULL");
    return false;
  }
  const int l[3] = {emf->ibeg, emf->jbeg, emf->kbeg}, h[3] = {emf->iend, emf->jend, emf->kend};
  for (int a = 0; a < 3; a++) {
    n[a] = grid->np_tot[a];
    lo[a] = l[a];
    hi[a] = h[a];
    if (lo[a] < 1 || hi[a] > n[a] - 3 || lo[a] > hi[a]) {
      report(fn, "EMF range outside [1, NX_TOT-3]");
      return false;
    }
  }
  return true;
}

}  // namespace

extern "C" void CT_Allocate(EMF* emf0)
{
  const long nx = NX1_TOT, ny = NX2_TOT, nz = NX3_TOT;
  memset(emf0, 0, sizeof(EMF));
  if (nx < 1 || ny < 1 || nz < 1) {
    report("CT_Allocate", "NX1_TOT, NX2_TOT, NX3_TOT are not set");
    return;
  }
  double**** all[] = {&emf0->ez, &emf0->ex, &emf0->ey, &emf0->ezi, &emf0->ezj, &emf0->exj,
                      &emf0->exk, &emf0->eyi, &emf0->eyk, &emf0->SxL, &emf0->SxR, &emf0->SyL,
                      &emf0->SyR, &emf0->SzL, &emf0->SzR, &emf0->dvx_dx, &emf0->dvx_dy,
                      &emf0->dvy_dx, &emf0->dvy_dy, &emf0->dvx_dz, &emf0->dvy_dz,
                      &emf0->dvz_dx, &emf0->dvz_dy, &emf0->dvz_dz, &emf0->dbx_dy, &emf0->dby_dx,
                      &emf0->dbx_dz, &emf0->dby_dz, &emf0->dbz_dx, &emf0->dbz_dy};
  for (double**** p : all) {
    *p = alloc3(nz, ny, nx);
    if (!*p) {
      report("CT_Allocate", "host allocation failed");
      return;
    }
  }
}

extern "C" void CT_StoreVelSlopes(EMF* emf, const Sweep* sweep, int beg, int end)
{
  const int i0 = g_i, j0 = g_j, k0 = g_k;
  double** vp = sweep->stateL.v;
  double** vm = sweep->stateR.v - 1;
  for (int q = beg; q <= end; q++) {
    const int i = g_dir == 0 ? q : i0, j = g_dir == 1 ? q : j0, k = g_dir == 2 ? q : k0;
    const double dx = vp[q][VX1] - vm[q][VX1];
    const double dy = vp[q][VX2] - vm[q][VX2];
    const double dz = vp[q][VX3] - vm[q][VX3];
    if (g_dir == 0) {
      emf->dvx_dx[k][j][i] = dx;
      emf->dvy_dx[k][j][i] = dy;
      emf->dvz_dx[k][j][i] = dz;
    } else if (g_dir == 1) {
      emf->dvx_dy[k][j][i] = dx;
      emf->dvy_dy[k][j][i] = dy;
      emf->dvz_dy[k][j][i] = dz;
    } else {
      emf->dvx_dz[k][j][i] = dx;
      emf->dvy_dz[k][j][i] = dy;
      emf->dvz_dz[k][j][i] = dz;
    }
  }
}

extern "C" void CT_StoreUpwindEMF(const Sweep* sweep, EMF* emf, int beg, int end, Grid* grid)
{
  (void)grid;
  const int i0 = g_i, j0 = g_j, k0 = g_k;
  for (int q = beg; q <= end; q++) {
    const int i = g_dir == 0 ? q : i0, j = g_dir == 1 ? q : j0, k = g_dir == 2 ? q : k0;
    const double* f = sweep->flux[q];
    const double sl = 0.0 >= -sweep->SL[q] ? 0.0 : -sweep->SL[q];   // MAX(0.0, -SL)
    const double sr = 0.0 >= sweep->SR[q] ? 0.0 : sweep->SR[q];     // MAX(0.0, SR)
    if (g_dir == 0) {
      emf->ezi[k][j][i] = -f[BX2];
      emf->eyi[k][j][i] = f[BX3];
      emf->SxL[k][j][i] = sl;
      emf->SxR[k][j][i] = sr;
    } else if (g_dir == 1) {
      emf->ezj[k][j][i] = f[BX1];
      emf->exj[k][j][i] = -f[BX3];
      emf->SyL[k][j][i] = sl;
      emf->SyR[k][j][i] = sr;
    } else {
      emf->eyk[k][j][i] = -f[BX1];
      emf->exk[k][j][i] = f[BX2];
      emf->SzL[k][j][i] = sl;
      emf->SzR[k][j][i] = sr;
    }
  }
  if (g_dir == 0) {
    emf->ibeg = beg;
    emf->iend = end;
  } else if (g_dir == 1) {
    emf->jbeg = beg;
    emf->jend = end;
  } else {
    emf->kbeg = beg;
    emf->kend = end;
  }
  CT_StoreVelSlopes(emf, sweep, beg, end + 1);
}

extern "C" void CT_ComputeEMF(const Data* d, Grid* grid)
{
  if (g_emf_average != PGPU_EMF_UCT_HLL) {
    report("CT_ComputeEMF", "the row-level CT entry points implement the UCT_HLL average");
    return;
  }
  EMF* emf = d->emf;
  int n[3], lo[3], hi[3];
  if (!ct_box(grid, emf, "CT_ComputeEMF", n, lo, hi)) return;
  const size_t N = (size_t)n[0] * n[1] * n[2];
  // the 21 inputs in EmfField order, then ex, ey, ez
  std::vector<double> h(24 * N);
  double*** in[21] = {emf->SxR, emf->SxL, emf->SyR, emf->SyL, emf->SzR, emf->SzL,
                      nullptr, nullptr, nullptr, d->Vc[VX1], d->Vc[VX2], d->Vc[VX3],
                      emf->dvx_dx, emf->dvy_dx, emf->dvz_dx, emf->dvx_dy, emf->dvy_dy,
                      emf->dvz_dy, emf->dvx_dz, emf->dvy_dz, emf->dvz_dz};
  for (int q = 0; q < 21; q++) {
    if (q >= 6 && q < 9) pack3(h.data() + q * N, d->Vs[q - 6], n);   // faces, i from 0
    else pack3(h.data() + q * N, in[q], n);
  }
  Scratch& S = scratch();
  double* dv = S.get(24 * N * sizeof(double));
  if (!dv) {
    report("CT_ComputeEMF", "device scratch allocation failed");
    return;
  }
  const double* f21[21];
  for (int q = 0; q < 21; q++) f21[q] = dv + q * N;
  bool ok = hipMemcpyAsync(dv, h.data(), 21 * N * sizeof(double), hipMemcpyHostToDevice, S.s) ==
            hipSuccess;
  ok = ok && launch_ct_emf_box(n, lo, hi, f21, dv + 21 * N, dv + 22 * N, dv + 23 * N, S.s) == 0;
  ok = ok && hipMemcpyAsync(h.data() + 21 * N, dv + 21 * N, 3 * N * sizeof(double),
                            hipMemcpyDeviceToHost, S.s) == hipSuccess;
  ok = ok && hipStreamSynchronize(S.s) == hipSuccess;
  if (!ok) {
    report("CT_ComputeEMF", "device launch failed");
    return;
  }
  unpack3_range(h.data() + 21 * N, emf->ex, n, lo, hi);
  unpack3_range(h.data() + 22 * N, emf->ey, n, lo, hi);
  unpack3_range(h.data() + 23 * N, emf->ez, n, lo, hi);
}

extern "C" void CT_Update(const Data* d, Data_Arr Bs, double dt, Grid* grid)
{
  EMF* emf = d->emf;
  int n[3], lo[3], hi[3];
  if (!ct_box(grid, emf, "CT_Update", n, lo, hi)) return;
  const size_t N = (size_t)n[0] * n[1] * n[2];
  const size_t W = (size_t)n[0] + n[1] + n[2];
  // ex, ey, ez, Bs (x, y, z), the updated faces (x, y, z; written over the
  // ranges only), then dx1, dx2, dx3
  std::vector<double> h(9 * N + W);
  pack3(h.data() + 0 * N, emf->ex, n);
  pack3(h.data() + 1 * N, emf->ey, n);
  pack3(h.data() + 2 * N, emf->ez, n);
  for (int a = 0; a < 3; a++) pack3(h.data() + (3 + a) * N, Bs[a], n);
  double* w = h.data() + 9 * N;
  memcpy(w, grid->dx[0], n[0] * sizeof(double));
  memcpy(w + n[0], grid->dx[1], n[1] * sizeof(double));
  memcpy(w + n[0] + n[1], grid->dx[2], n[2] * sizeof(double));
  Scratch& S = scratch();
  double* dv = S.get((9 * N + W) * sizeof(double));
  if (!dv) {
    report("CT_Update", "device scratch allocation failed");
    return;
  }
  const double* b0[3] = {dv + 3 * N, dv + 4 * N, dv + 5 * N};
  double* b[3] = {dv + 6 * N, dv + 7 * N, dv + 8 * N};
  const double* dw = dv + 9 * N;
  bool ok = hipMemcpyAsync(dv, h.data(), 6 * N * sizeof(double), hipMemcpyHostToDevice, S.s) ==
            hipSuccess;
  ok = ok && hipMemcpyAsync(dv + 9 * N, w, W * sizeof(double), hipMemcpyHostToDevice, S.s) ==
                 hipSuccess;
  ok = ok && launch_ct_update_box(n, lo, hi, dt, dw, dw + n[0], dw + n[0] + n[1], dv, dv + N,
                                  dv + 2 * N, b0, b, S.s) == 0;
  ok = ok && hipMemcpyAsync(h.data() + 6 * N, dv + 6 * N, 3 * N * sizeof(double),
                            hipMemcpyDeviceToHost, S.s) == hipSuccess;
  ok = ok && hipStreamSynchronize(S.s) == hipSuccess;
  if (!ok) {
    report("CT_Update", "device launch failed");
    return;
  }
  // each component over the range the reference updates (ct_update.c:79-190)
  const int lx[3] = {lo[0], lo[1] + 1, lo[2] + 1}, ly[3] = {lo[0] + 1, lo[1], lo[2] + 1};
  const int lz[3] = {lo[0] + 1, lo[1] + 1, lo[2]};
  unpack3_range(h.data() + 6 * N, d->Vs[0], n, lx, hi);
  unpack3_range(h.data() + 7 * N, d->Vs[1], n, ly, hi);
  unpack3_range(h.data() + 8 * N, d->Vs[2], n, lz, hi);
}


// ---------------------------------------------------------------------
// Boundary (boundary.c:40-315) on PLUTO's ghosted arrays: the ghost zones
// of Vc and Vs filled on the device, side by side in the reference's order.
// ---------------------------------------------------------------------

namespace {

// a box array whose axis `ext` (0..2, or -1) starts at index -1, as
// [k][j][i] rows <-> a device-layout block
void box_copy(double* h, double*** a, const GhostBox& b, int ext, bool to_host)
{
  const int l0 = ext == 0 ? -1 : 0, l1 = ext == 1 ? -1 : 0, l2 = ext == 2 ? -1 : 0;
  const size_t row = (size_t)(b.t[0] - l0);
  size_t o = 0;
  for (int k = l2; k < b.t[2]; k++)
    for (int j = l1; j < b.t[1]; j++, o += row) {
      if (to_host) memcpy(&a[k][j][l0], h + o, row * sizeof(double));
      else memcpy(h + o, &a[k][j][l0], row * sizeof(double));
    }
}

size_t box_size(const GhostBox& b, int ext)
{
  size_t s = 1;
  for (int a = 0; a < 3; a++) s *= (size_t)(b.t[a] + (a == ext));
  return s;
}

}  // namespace

extern "C" void Boundary(const Data* d, int idim, Grid* grid)
{
  const char* fn = "Boundary";
  if (grid->np_int[2] <= 1) {
    report(fn, "the PLUTO entry points support DIMENSIONS 3 (COMPONENTS 3) builds");
    return;
  }
  if (idim < -1 || idim > 2) {
    report(fn, "idim must be IDIR, JDIR, KDIR or ALL_DIR");
    return;
  }
  GhostBox b;
  int type[6];
  bool need_area = false;
  for (int a = 0; a < 3; a++) {
    b.t[a] = grid->np_tot[a];
    b.beg[a] = grid->lbeg[a];
    b.end[a] = grid->lend[a];
    b.n[a] = b.end[a] - b.beg[a] + 1;
    if (b.beg[a] < 1 || b.end[a] > b.t[a] - 2 || b.n[a] != grid->np_int[a] ||
        b.n[a] < b.beg[a] || b.n[a] < b.t[a] - 1 - b.end[a]) {
      report(fn, "unexpected grid extents (one ghost zone at least per side, and no more "
                 "ghost zones than interior cells)");
      return;
    }
    if (grid->nproc[a] > 1) {
      report(fn, "a domain split over processes is not supported");
      return;
    }
    type[2 * a] = grid->lbound[a];
    type[2 * a + 1] = grid->rbound[a];
  }
  const int sbeg = idim < 0 ? 0 : 2 * idim, send = idim < 0 ? 5 : 2 * idim + 1;
  for (int is = sbeg; is <= send; is++) {
    const int ty = type[is];
    if (ty == 0 || ty == BOUND_PERIODIC) continue;
    if (ty == BOUND_USERDEF && !UserDefBoundary) {
      report(fn, "a USERDEF side needs the program's UserDefBoundary");
      return;
    }
    if ((ty < BOUND_OUTFLOW || ty > BOUND_EQTSYMMETRIC) && ty != BOUND_USERDEF) {
      report(fn, "shearing-box boundaries are not supported on the device");
      return;
    }
    need_area = true;
  }
  if (need_area && (!grid->A[0] || !grid->A[1] || !grid->A[2])) {
    report(fn, "grid->A is NULL");
    return;
  }
  // device blocks: Vc (8), Vs (3), then the areas (3) when a side needs them
  size_t off[14], tot = 0;
  for (int q = 0; q < 14; q++) {
    off[q] = tot;
    const int ext = q < 8 ? -1 : (q - 8) % 3;
    if (q < 11 || need_area) tot += box_size(b, ext);
  }
  std::vector<double> h(tot);
  for (int q = 0; q < 8; q++) box_copy(h.data() + off[q], d->Vc[q], b, -1, false);
  for (int a = 0; a < 3; a++) box_copy(h.data() + off[8 + a], d->Vs[a], b, a, false);
  if (need_area)
    for (int a = 0; a < 3; a++) box_copy(h.data() + off[11 + a], grid->A[a], b, a, false);
  Scratch& S = scratch();
  double* dv = S.get(tot * sizeof(double));
  if (!dv) {
    report(fn, "device scratch allocation failed");
    return;
  }
  double* vc[8];
  double* vs[3];
  const double* area[3];
  for (int q = 0; q < 8; q++) vc[q] = dv + off[q];
  for (int a = 0; a < 3; a++) {
    vs[a] = dv + off[8 + a];
    area[a] = need_area ? dv + off[11 + a] : nullptr;
  }
  bool ok = hipMemcpyAsync(dv, h.data(), tot * sizeof(double), hipMemcpyHostToDevice, S.s) ==
            hipSuccess;
  for (int is = sbeg; is <= send && ok; is++) {
    if (type[is] == 0) continue;
    if (type[is] == BOUND_USERDEF) {
      // the sides before it are done: the hook sees PLUTO's arrays as they
      // stand, then FillMagneticField and the normal average run on the device
      ok = hipMemcpyAsync(h.data(), dv, off[11] * sizeof(double), hipMemcpyDeviceToHost, S.s) ==
               hipSuccess &&
           hipStreamSynchronize(S.s) == hipSuccess;
      if (!ok) break;
      for (int q = 0; q < 8; q++) box_copy(h.data() + off[q], d->Vc[q], b, -1, true);
      for (int a = 0; a < 3; a++) box_copy(h.data() + off[8 + a], d->Vs[a], b, a, true);
      userdef_on_host(d, grid, b.t, b.beg, b.end, is);
      for (int q = 0; q < 8; q++) box_copy(h.data() + off[q], d->Vc[q], b, -1, false);
      for (int a = 0; a < 3; a++) box_copy(h.data() + off[8 + a], d->Vs[a], b, a, false);
      ok = hipMemcpyAsync(dv, h.data(), off[11] * sizeof(double), hipMemcpyHostToDevice, S.s) ==
           hipSuccess;
      ok = ok && launch_boundary_fill_normal(b, is, vc, vs, area, nullptr, S.s) == 0;
      continue;
    }
    ok = launch_boundary_side(b, is, type[is], vc, vs, area, nullptr, S.s) == 0;
  }
  ok = ok && hipMemcpyAsync(h.data(), dv, off[11] * sizeof(double), hipMemcpyDeviceToHost,
                            S.s) == hipSuccess;
  ok = ok && hipStreamSynchronize(S.s) == hipSuccess;
  if (!ok) {
    report(fn, "device launch failed");
    return;
  }
  for (int q = 0; q < 8; q++) box_copy(h.data() + off[q], d->Vc[q], b, -1, true);
  for (int a = 0; a < 3; a++) box_copy(h.data() + off[8 + a], d->Vs[a], b, a, true);
}
