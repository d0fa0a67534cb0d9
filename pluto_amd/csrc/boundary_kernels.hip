// Ghost-zone filling of PLUTO's ghosted boxes on the device: one side of
// Boundary() (boundary.c:120-292) per launch pair.
//
//   k_bound_fill    the cell-centred primitives and the staggered field over
//                   the side's boxes (boundary.c:144-165): OutflowBound,
//                   PeriodicBound, ReflectiveBound with FlipSign's signs
//                   (boundary.c:318-564).  One thread per transverse point
//                   of one array, marching the normal index in the order of
//                   BOX_LOOP (macros.h:91-94), so a source inside the same
//                   box is read exactly when the reference reads it.
//   k_bound_divb    FillMagneticField (MHD/CT/ct_fill_mag_field.c:38-140):
//                   the normal staggered field of the ghost faces from
//                   div.B = 0, marching outward from the domain; for outflow
//                   sides also CT_AverageNormalMagField
//                   (MHD/CT/ct_field_average.c:134-270, Cartesian) along the
//                   same line.
//
// Lines along the normal are independent: a transverse neighbour read by a
// line is a tangential component or an area, which neither kernel writes
// for that side.

#include <hip/hip_runtime.h>
#include "boundary_kernels.hpp"

namespace pgpu {

namespace {

constexpr int kMaxOps = 11;     // 8 primitives + 3 staggered components
constexpr int kVX1 = 1, kBX1 = 4;

// element (k, j, i) of a box array whose axis `ext` (0..2, or -1 for none)
// holds the extra face plane -1
__device__ __host__ inline long gidx(const GhostBox& b, int ext, int i, int j, int k)
{
  if (b.uni) return ((long)(k + 1) * (b.t[1] + 1) + (j + 1)) * (b.t[0] + 1) + (i + 1);
  const long n0 = b.t[0] + (ext == 0), n1 = b.t[1] + (ext == 1);
  return ((long)(k + (ext == 2)) * n1 + (j + (ext == 1))) * n0 + (i + (ext == 0));
}

struct FillOp {
  double* p;
  int ext;          // axis of the face plane -1, or -1
  int lo[3], hi[3]; // RBox beg/end per axis (the normal one may descend)
  double sgn;       // ReflectiveBound's sign (1 otherwise)
};

struct FillSide {
  FillOp op[kMaxOps];
  int nop, dir, end_side, kind;   // kind: 0 outflow, 1 periodic, 2 reflective
};

__global__ void __launch_bounds__(256) k_bound_fill(GhostBox b, FillSide f)
{
  const FillOp& o = f.op[blockIdx.y];
  const int d = f.dir, ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
  const int na = o.hi[ta] - o.lo[ta] + 1, nb = o.hi[tb] - o.lo[tb] + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)na * nb) return;
  int p[3];
  p[ta] = o.lo[ta] + (int)(t % na);
  p[tb] = o.lo[tb] + (int)(t / na);
  const int step = o.lo[d] <= o.hi[d] ? 1 : -1;
  const int N = b.n[d];
  for (int m = o.lo[d]; m != o.hi[d] + step; m += step) {
    int q[3] = {p[0], p[1], p[2]};
    q[d] = f.kind == 0 ? (f.end_side ? b.end[d] : b.beg[d])
         : f.kind == 1 ? (f.end_side ? m - N : m + N)
                       : (f.end_side ? 2 * b.end[d] - m + 1 : 2 * b.beg[d] - m - 1);
    p[d] = m;
    const double v = o.p[gidx(b, o.ext, q[0], q[1], q[2])];
    o.p[gidx(b, o.ext, p[0], p[1], p[2])] = f.kind == 2 ? o.sgn * v : v;
  }
}

struct DivbSide {
  double* bs[3];
  const double* area[3];  // or null: area_c
  double area_c[3];
  double* bc;           // the cell-centred normal field, or null (no average)
  int dir, end_side;
};

__global__ void __launch_bounds__(256) k_bound_divb(GhostBox b, DivbSide f)
{
  const int d = f.dir, ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
  const int na = b.t[ta], nb = b.t[tb];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)na * nb) return;
  int p[3];
  p[ta] = (int)(t % na);
  p[tb] = (int)(t / na);
  double* bn = f.bs[d];
  auto A = [&](int e, long idx) { return f.area[e] ? f.area[e][idx] : f.area_c[e]; };
  // the transverse terms A(+) b(+) - A(-) b(-) of div.B at cell p, in axis order
  auto dB = [&](int e) {
    int m[3] = {p[0], p[1], p[2]};
    m[e] -= 1;
    const long ip = gidx(b, e, p[0], p[1], p[2]), im = gidx(b, e, m[0], m[1], m[2]);
    return A(e, ip) * f.bs[e][ip] - A(e, im) * f.bs[e][im];
  };
  if (!f.end_side) {
    for (int m = b.beg[d] - 1; m >= 0; m--) {
      p[d] = m;
      int q[3] = {p[0], p[1], p[2]};
      q[d] = m - 1;
      const long ip = gidx(b, d, p[0], p[1], p[2]), im = gidx(b, d, q[0], q[1], q[2]);
      const double d1 = dB(ta), d2 = dB(tb);
      bn[im] = (A(d, ip) * bn[ip] + d1 + d2) / A(d, im);
    }
  } else {
    for (int m = b.end[d] + 1; m < b.t[d]; m++) {
      p[d] = m;
      int q[3] = {p[0], p[1], p[2]};
      q[d] = m - 1;
      const long ip = gidx(b, d, p[0], p[1], p[2]), im = gidx(b, d, q[0], q[1], q[2]);
      const double d1 = dB(ta), d2 = dB(tb);
      bn[ip] = (A(d, im) * bn[im] - (d1 + d2)) / A(d, ip);
    }
  }
  if (!f.bc) return;
  const int lo = f.end_side ? b.end[d] + 1 : 0, hi = f.end_side ? b.t[d] - 1 : b.beg[d] - 1;
  for (int m = lo; m <= hi; m++) {
    p[d] = m;
    int q[3] = {p[0], p[1], p[2]};
    q[d] = m - 1;
    f.bc[gidx(b, -1, p[0], p[1], p[2])] =
        0.5 * (bn[gidx(b, d, p[0], p[1], p[2])] + bn[gidx(b, d, q[0], q[1], q[2])]);
  }
}

// FlipSign (boundary.c:318-430) for MHD with three components
void flip_sign(int dir, int type, double sgn[8])
{
  for (int q = 0; q < 8; q++) sgn[q] = 1.0;
  const int n = dir, p = dir == 0 ? 1 : 0, t = dir == 2 ? 1 : 2;
  sgn[kVX1 + n] = -1.0;
  sgn[kBX1 + n] = -1.0;
  if (type == BOUND_AXISYMMETRIC) {
    sgn[kVX1 + t] = -1.0;
    sgn[kBX1 + t] = -1.0;
  }
  if (type == BOUND_EQTSYMMETRIC) {
    sgn[kBX1 + n] = 1.0;
    sgn[kBX1 + p] = -1.0;
    sgn[kBX1 + t] = -1.0;
  }
}

}  // namespace

// FillMagneticField on side (d, end_side) and, with bc set, the
// cell-centred normal field of CT_AverageNormalMagField
static int launch_divb_side(const GhostBox& b, int d, int end_side, double* bc, double* const* vs,
                            const double* const* area, const double* area_c, hipStream_t s)
{
  DivbSide g{};
  for (int a = 0; a < 3; a++) {
    g.bs[a] = vs[a];
    g.area[a] = area ? area[a] : nullptr;
    g.area_c[a] = area_c ? area_c[a] : 0.0;
  }
  if (!area && !area_c) return (int)hipErrorInvalidValue;
  g.bc = bc;
  g.dir = d;
  g.end_side = end_side;
  const long lines = (long)b.t[d == 0 ? 1 : 0] * b.t[d == 2 ? 1 : 2];
  hipLaunchKernelGGL(k_bound_divb, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, s, b, g);
  return (int)hipGetLastError();
}

int launch_boundary_side(const GhostBox& b, int side, int type, double* const* vc,
                         double* const* vs, const double* const* area, const double* area_c,
                         hipStream_t s)
{
  if (side < 0 || side > 5) return (int)hipErrorInvalidValue;
  const int d = side / 2, end_side = side & 1;
  FillSide f{};
  f.dir = d;
  f.end_side = end_side;
  if (type == BOUND_OUTFLOW) f.kind = 0;
  else if (type == BOUND_PERIODIC) f.kind = 1;
  else if (type == BOUND_REFLECTIVE || type == BOUND_AXISYMMETRIC || type == BOUND_EQTSYMMETRIC)
    f.kind = 2;
  else return (int)hipErrorInvalidValue;
  double sgn[8];
  flip_sign(d, type, sgn);
  // the centre box (boundary.c:144-154)
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = 0;
    hi[a] = b.t[a] - 1;
  }
  lo[d] = end_side ? b.end[d] + 1 : b.beg[d] - 1;
  hi[d] = end_side ? b.t[d] - 1 : 0;
  long most = 0;
  auto add = [&](double* p, int ext, const int* l, const int* h, double sg) {
    FillOp& o = f.op[f.nop++];
    o.p = p;
    o.ext = ext;
    for (int a = 0; a < 3; a++) {
      o.lo[a] = l[a];
      o.hi[a] = h[a];
    }
    o.sgn = sg;
    long c = 1;
    for (int a = 0; a < 3; a++)
      if (a != d) c *= (long)(h[a] - l[a] + 1);
    if (c > most) most = c;
  };
  for (int q = 0; q < 8; q++) add(vc[q], -1, lo, hi, sgn[q]);
  // the face boxes (boundary.c:157-165): one more face below along their own
  // axis; the normal component only for periodic sides
  for (int a = 0; a < 3; a++) {
    if (a == d && f.kind != 1) continue;
    int l[3] = {lo[0], lo[1], lo[2]}, h[3] = {hi[0], hi[1], hi[2]};
    if (l[a] < h[a]) l[a] -= 1;
    else h[a] -= 1;
    add(vs[a], a, l, h, sgn[kBX1 + a]);
  }
  const dim3 bl(256);
  hipLaunchKernelGGL(k_bound_fill, dim3((unsigned)((most + 255) / 256), f.nop), bl, 0, s, b, f);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || f.kind == 1) return (int)e;
  return launch_divb_side(b, d, end_side, f.kind == 0 ? vc[kBX1 + d] : nullptr, vs, area, area_c,
                          s);
}

int launch_boundary_fill_normal(const GhostBox& b, int side, double* const* vc, double* const* vs,
                                const double* const* area, const double* area_c, hipStream_t s)
{
  if (side < 0 || side > 5) return (int)hipErrorInvalidValue;
  const int d = side / 2;
  return launch_divb_side(b, d, side & 1, vc[kBX1 + d], vs, area, area_c, s);
}

}  // namespace pgpu
