// Discrete adjoint (vector-Jacobian products) of the point functions of the
// MHD stage, and a forward-mode dual number that runs the precision-generic
// forward templates of mhd_device.hpp as their tangent-linear model.
//
// Each *_vjp function recomputes its forward values from the same inputs and
// then accumulates lambda^T (d out / d in) into the a_* adjoint arrays
// (+=).  Non-smooth operations take the derivative of the branch the
// forward pass took: the MC limiter its active argument, min/max their
// selected operand, HLLD its region (Miyoshi & Kusano 2005; MHD/hlld.c:44-452)
// and its degeneracy switch.  The tangent model (Dual) takes the same
// branches, so <lambda, J t> = <J^T lambda, t> holds to round-off; the tests
// check exactly that.

#pragma once
#include <hip/hip_runtime.h>
#include "mhd_device.hpp"

namespace pgpu {

// ---------------------------------------------------------------------
// forward-mode dual number
// ---------------------------------------------------------------------

// (in its own namespace: its sqrt/fabs overloads are found by argument-
// dependent lookup from the templates without hiding the double versions)
namespace ad {

struct Dual {
  double v, d;
  Dual() = default;
  __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
  __host__ __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
};

__host__ __device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator*(Dual a, Dual b)
{
  return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
__host__ __device__ inline Dual operator/(Dual a, Dual b)
{
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__host__ __device__ inline Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
__host__ __device__ inline Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
__host__ __device__ inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
__host__ __device__ inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
__host__ __device__ inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__host__ __device__ inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__host__ __device__ inline Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
__host__ __device__ inline Dual sqrt(Dual a)
{
  const double s = ::sqrt(a.v);
  return Dual(s, a.d / (2.0 * s));
}
}  // namespace ad

using ad::Dual;
template <> __device__ __forceinline__ Dual div_(Dual a, Dual b) { return a / b; }
template <> __device__ __forceinline__ Dual sqrt_(Dual x) { return sqrt(x); }

// ---------------------------------------------------------------------
// point VJPs
// ---------------------------------------------------------------------

// The reverse passes below may fuse multiply-adds: their values are
// derivatives, checked against finite differences, not bitwise against the
// reference.  (The forward functions they call, and Dual above, keep the
// library-wide no-contraction rule, so recomputed forward values and the
// tangent's value part stay exact.)
#pragma clang fp contract(fast)

// s = MC slope(vm1, v0, vp1) (plm_coeffs.h:119-123): adjoint of s.
__device__ __forceinline__ void mc_slope_vjp(double vm1, double v0, double vp1, double a_s,
                                             double& a_vm1, double& a_v0, double& a_vp1)
{
  const double dvp = vp1 - v0;
  const double dvm = v0 - vm1;
  if (!(dvp * dvm > 0.0)) return;
  const double qc = 0.5 * (dvm + dvp);
  const bool p_small = fabs(dvp) < fabs(dvm);
  const double sc = 2.0 * (p_small ? dvp : dvm);
  double a_dvp = 0.0, a_dvm = 0.0;
  if (fabs(qc) < fabs(sc)) {
    a_dvp = 0.5 * a_s;
    a_dvm = 0.5 * a_s;
  } else if (p_small) {
    a_dvp = 2.0 * a_s;
  } else {
    a_dvm = 2.0 * a_s;
  }
  a_vp1 += a_dvp;
  a_v0 += a_dvm - a_dvp;
  a_vm1 -= a_dvm;
}

// u = PrimToCons(v) (mappers.c:25-86): a_v += (du/dv)^T a_u.
__device__ __forceinline__ void prim2cons_vjp(const double* v, const double* a_u, double* a_v,
                                              double gmm1)
{
  const double v2 = v[VX1] * v[VX1] + v[VX2] * v[VX2] + v[VX3] * v[VX3];
  const double aE = a_u[ENG];
  a_v[RHO] += a_u[RHO] + a_u[VX1] * v[VX1] + a_u[VX2] * v[VX2] + a_u[VX3] * v[VX3] + 0.5 * v2 * aE;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    a_v[VX1 + k] += v[RHO] * (a_u[VX1 + k] + aE * v[VX1 + k]);
    a_v[BX1 + k] += a_u[BX1 + k] + aE * v[BX1 + k];
  }
  a_v[PRS] += aE / gmm1;
}

// v = ConsToPrim(u) without floors (mappers.c:140-189), written in terms of
// the primitive result v: a_u += (dv/du)^T a_v over rho, m, E; the field
// part of the pressure adjoint goes to a_b (the cell-centred field).
__device__ __forceinline__ void cons2prim_vjp(const double* v, const double* a_v, double* a_u,
                                              double* a_b, double gmm1)
{
  const double rho = v[RHO], tau = 1.0 / rho;
  const double ap = a_v[PRS];
  const double v2 = v[VX1] * v[VX1] + v[VX2] * v[VX2] + v[VX3] * v[VX3];
  double a_rho = a_v[RHO] + 0.5 * gmm1 * v2 * ap;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    a_rho -= a_v[VX1 + k] * v[VX1 + k] * tau;
    a_u[1 + k] += a_v[VX1 + k] * tau - gmm1 * v[VX1 + k] * ap;
    a_b[k] += -gmm1 * v[BX1 + k] * ap;
  }
  a_u[0] += a_rho;
  a_u[4] += gmm1 * ap;
}

// Normal flux (fluxes.c:159-215) with u = PrimToCons(v): adjoint into a_v
// (directly) and a_u (to be chained through prim2cons_vjp by the caller).
template <int D>
__device__ __forceinline__ void normal_flux_vjp(const double* v, const double* u,
                                                const double* a_f, double a_pt, double* a_v,
                                                double* a_u)
{
  constexpr int VN = VX1 + Dir<D>::n, VT = VX1 + Dir<D>::t, VB = VX1 + Dir<D>::b;
  constexpr int BN = BX1 + Dir<D>::n, BT = BX1 + Dir<D>::t, BB = BX1 + Dir<D>::b;
  const double bmag2 = v[BX1] * v[BX1] + v[BX2] * v[BX2] + v[BX3] * v[BX3];
  const double pt = v[PRS] + 0.5 * bmag2;
  const double vb = v[VX1] * v[BX1] + v[VX2] * v[BX2] + v[VX3] * v[BX3];
  // fx[RHO] = u[VN]
  a_u[VN] += a_f[RHO];
  // fx[VXk] = v[VN] u[VXk] - v[BN] v[BXk]
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double a = a_f[VX1 + k];
    a_v[VN] += a * u[VX1 + k];
    a_u[VX1 + k] += a * v[VN];
    a_v[BN] -= a * v[BX1 + k];
    a_v[BX1 + k] -= a * v[BN];
  }
  // fx[BT] = v[VN] v[BT] - v[BN] v[VT]; fx[BB] likewise; fx[BN] = 0
  {
    const double a = a_f[BT];
    a_v[VN] += a * v[BT];
    a_v[BT] += a * v[VN];
    a_v[BN] -= a * v[VT];
    a_v[VT] -= a * v[BN];
  }
  {
    const double a = a_f[BB];
    a_v[VN] += a * v[BB];
    a_v[BB] += a * v[VN];
    a_v[BN] -= a * v[VB];
    a_v[VB] -= a * v[BN];
  }
  // fx[ENG] = (u[ENG] + pt) v[VN] - v[BN] vb
  {
    const double a = a_f[ENG];
    a_u[ENG] += a * v[VN];
    a_pt += a * v[VN];
    a_v[VN] += a * (u[ENG] + pt);
    a_v[BN] -= a * vb;
    const double a_vb = -a * v[BN];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      a_v[VX1 + k] += a_vb * v[BX1 + k];
      a_v[BX1 + k] += a_vb * v[VX1 + k];
    }
  }
  // pt = p + 0.5 |B|^2
  a_v[PRS] += a_pt;
#pragma unroll
  for (int k = 0; k < 3; k++) a_v[BX1 + k] += a_pt * v[BX1 + k];
}

// cmin = v_n - cf, cmax = v_n + cf (eigenv.c:35-104, hll_speed.c:76-107).
template <int D>
__device__ __forceinline__ void fast_speeds_vjp(const double* v, double gamma, double a_cmin,
                                                double a_cmax, double* a_v)
{
  constexpr int VN = VX1 + Dir<D>::n;
  constexpr int BN = BX1 + Dir<D>::n, BT = BX1 + Dir<D>::t, BB = BX1 + Dir<D>::b;
  const double gpr = gamma * v[PRS];
  const double b1 = v[BN], b2 = v[BT], b3 = v[BB];
  const double btmag2 = b2 * b2 + b3 * b3;
  const double bmag2 = b1 * b1 + btmag2;
  const double cf0 = gpr - bmag2;
  const double q = cf0 * cf0 + 4.0 * gpr * btmag2;
  const double sq = sqrt(q);
  const double cf1 = gpr + bmag2 + sq;
  const double w = 0.5 * cf1 / v[RHO];
  const double cf = sqrt(w);
  a_v[VN] += a_cmin + a_cmax;
  const double a_cf = a_cmax - a_cmin;
  const double a_w = a_cf / (2.0 * cf);
  const double a_cf1 = a_w * 0.5 / v[RHO];
  a_v[RHO] -= a_w * w / v[RHO];
  double a_gpr = a_cf1, a_bmag2 = a_cf1;
  const double a_q = sq > 0.0 ? a_cf1 / (2.0 * sq) : 0.0;
  const double a_cf0 = a_q * 2.0 * cf0;
  a_gpr += a_q * 4.0 * btmag2 + a_cf0;
  double a_btmag2 = a_q * 4.0 * gpr;
  a_bmag2 -= a_cf0;
  a_v[BN] += 2.0 * b1 * a_bmag2;
  a_btmag2 += a_bmag2;
  a_v[BT] += 2.0 * b2 * a_btmag2;
  a_v[BB] += 2.0 * b3 * a_btmag2;
  a_v[PRS] += gamma * a_gpr;
}

// Adjoint seeds of one HLLD interface: the flux components the stage uses
// (rho, m_x, m_y, m_z, E), the total pressure, and the Davis speeds SL, SR.
struct HlldSeed {
  double f[NV];   // only RHO, VX1..VX3, ENG are read
  double press, SL, SR;
};

// HLLD (non-FULL outputs) VJP: a_vL, a_vR += J^T seed.  vL/vR hold the
// face field in their normal slot (separate inputs here; the caller adds
// the two normal-slot adjoints to the face).
template <int D>
__device__ void hlld_vjp(const double* vL, const double* vR, double gamma, double gmm1,
                         const HlldSeed& s, double* a_vL, double* a_vR)
{
  constexpr int VN = VX1 + Dir<D>::n, VT = VX1 + Dir<D>::t, VB = VX1 + Dir<D>::b;
  constexpr int BN = BX1 + Dir<D>::n, BT = BX1 + Dir<D>::t, BB = BX1 + Dir<D>::b;
  constexpr int OUT[5] = {RHO, VX1, VX2, VX3, ENG};

  // ---------------- forward recomputation ----------------
  double uL[NV], uR[NV], fL[NV], fR[NV], ptL, ptR;
  prim2cons(vL, uL, gmm1);
  prim2cons(vR, uR, gmm1);
  normal_flux<D, true>(vL, uL, fL, ptL);
  normal_flux<D, true>(vR, uR, fR, ptR);
  double slmin, slmax, srmin, srmax;
  fast_speeds<D>(vL, gamma, slmin, slmax);
  fast_speeds<D>(vR, gamma, srmin, srmax);
  const bool selL = slmin <= srmin;     // SL = min_le(slmin, srmin)
  const bool selR = slmax >= srmax;     // SR = max_ge(slmax, srmax)
  const double SL = selL ? slmin : srmin;
  const double SR = selR ? slmax : srmax;

  double a_uL[NV] = {0}, a_uR[NV] = {0}, a_fL[NV] = {0}, a_fR[NV] = {0};
  double a_ptL = 0.0, a_ptR = 0.0, a_SL = s.SL, a_SR = s.SR;

  if (SL >= 0.0) {
#pragma unroll
    for (int q = 0; q < 5; q++) a_fL[OUT[q]] += s.f[OUT[q]];
    a_ptL += s.press;
  } else if (SR <= 0.0) {
#pragma unroll
    for (int q = 0; q < 5; q++) a_fR[OUT[q]] += s.f[OUT[q]];
    a_ptR += s.press;
  } else {
    // F1-F9
    const double t1 = SR - SL;
    const double scrh1 = 1.0 / t1;
    const double num2 = SR * vR[BN] - SL * vL[BN];
    const double Bx1 = num2 * scrh1;
    const double Bx = Bx1;
    const double sBx = (Bx > 0.0 ? 1.0 : -1.0);
    const double duL = SL - vL[VN];
    const double duR = SR - vR[VN];
    const double den2 = duR * uR[RHO] - duL * uL[RHO];
    const double scrh2 = 1.0 / den2;
    const double num5 = duR * uR[VN] - duL * uL[VN] - ptR + ptL;
    const double SM = num5 * scrh2;
    const double dvn = vR[VN] - vL[VN];
    const double rr = vL[RHO] * vR[RHO];
    const double num6 = duR * uR[RHO] * ptL - duL * uL[RHO] * ptR + rr * duR * duL * dvn;
    const double pts = num6 * scrh2;
    const double dL = SL - SM, dR = SR - SM;
    double usL[NV], usR[NV];
    usL[RHO] = uL[RHO] * duL / dL;
    usR[RHO] = uR[RHO] * duR / dR;
    const double sqrL = sqrt(usL[RHO]);
    const double sqrR = sqrt(usR[RHO]);
    double S1L = SM - fabs(Bx) / sqrL;
    double S1R = SM + fabs(Bx) / sqrR;
    bool revert = false;
    if ((S1L - SL) < 1.e-4 * (SM - SL)) revert = true;
    if ((S1R - SR) > -1.e-4 * (SR - SM)) revert = true;
    // F10
    double scrh3 = 0.0, nbn = 0.0, nbt = 0.0, nbb = 0.0;
    double sLn = 0, sLd = 0, sL = 0, sRn = 0, sRd = 0, sR = 0;
    if (revert) {
      scrh3 = 1.0 / (SR - SL);
      nbn = SR * uR[BN] - SL * uL[BN] + fL[BN] - fR[BN];
      nbt = SR * uR[BT] - SL * uL[BT] + fL[BT] - fR[BT];
      nbb = SR * uR[BB] - SL * uL[BB] + fL[BB] - fR[BB];
      usL[BN] = usR[BN] = nbn * scrh3;
      usL[BT] = usR[BT] = nbt * scrh3;
      usL[BB] = usR[BB] = nbb * scrh3;
      S1L = S1R = SM;
    } else {
      sLn = uL[RHO] * duL * duL - Bx * Bx;
      sLd = uL[RHO] * duL * dL - Bx * Bx;
      sL = sLn / sLd;
      sRn = uR[RHO] * duR * duR - Bx * Bx;
      sRd = uR[RHO] * duR * dR - Bx * Bx;
      sR = sRn / sRd;
      usL[BN] = Bx1;
      usL[BT] = uL[BT] * sL;
      usL[BB] = uL[BB] * sL;
      usR[BN] = Bx1;
      usR[BT] = uR[BT] * sR;
      usR[BB] = uR[BB] * sR;
    }
    // F11-F14
    const double ddL = uL[RHO] * duL, ddR = uR[RHO] * duR;
    const double scrhL = Bx / ddL, scrhR = Bx / ddR;
    const double vsL = vL[VT] - scrhL * (usL[BT] - uL[BT]);
    const double vsR = vR[VT] - scrhR * (usR[BT] - uR[BT]);
    const double wsL = vL[VB] - scrhL * (usL[BB] - uL[BB]);
    const double wsR = vR[VB] - scrhR * (usR[BB] - uR[BB]);
    usL[VN] = usL[RHO] * SM;
    usR[VN] = usR[RHO] * SM;
    usL[VT] = usL[RHO] * vsL;
    usR[VT] = usR[RHO] * vsR;
    usL[VB] = usL[RHO] * wsL;
    usR[VB] = usR[RHO] * wsR;
    const double cL = (vL[VN] * Bx1 + vL[VT] * uL[BT] + vL[VB] * uL[BB]) -
                      (SM * Bx1 + vsL * usL[BT] + wsL * usL[BB]);
    const double cR = (vR[VN] * Bx1 + vR[VT] * uR[BT] + vR[VB] * uR[BB]) -
                      (SM * Bx1 + vsR * usR[BT] + wsR * usR[BB]);
    const double nEL = duL * uL[ENG] - ptL * vL[VN] + pts * SM + Bx * cL;
    const double nER = duR * uR[ENG] - ptR * vR[VN] + pts * SM + Bx * cR;
    usL[ENG] = nEL / dL;
    usR[ENG] = nER / dR;

    double a_usL[NV] = {0}, a_usR[NV] = {0};
    double a_S1L = 0.0, a_S1R = 0.0, a_SM = 0.0;
    double a_vsL = 0.0, a_vsR = 0.0, a_wsL = 0.0, a_wsR = 0.0;
    double a_sqrL = 0.0, a_sqrR = 0.0, a_Bx = 0.0;

    if (S1L >= 0.0) {
#pragma unroll
      for (int q = 0; q < 5; q++) {
        const int nv = OUT[q];
        const double a = s.f[nv];
        a_fL[nv] += a;
        a_SL += a * (usL[nv] - uL[nv]);
        a_usL[nv] += a * SL;
        a_uL[nv] -= a * SL;
      }
      a_ptL += s.press;
    } else if (S1R <= 0.0) {
#pragma unroll
      for (int q = 0; q < 5; q++) {
        const int nv = OUT[q];
        const double a = s.f[nv];
        a_fR[nv] += a;
        a_SR += a * (usR[nv] - uR[nv]);
        a_usR[nv] += a * SR;
        a_uR[nv] -= a * SR;
      }
      a_ptR += s.press;
    } else {
      // F15-F18 forward
      const double den = sqrL + sqrR;
      const double nvss = sqrL * vsL + sqrR * vsR + (usR[BT] - usL[BT]) * sBx;
      const double vss = nvss / den;
      const double nwss = sqrL * wsL + sqrR * wsR + (usR[BB] - usL[BB]) * sBx;
      const double wss = nwss / den;
      double uss[NV], ussl[NV], ussr[NV];
      ussl[RHO] = usL[RHO];
      ussr[RHO] = usR[RHO];
      ussl[VN] = ussl[RHO] * SM;
      ussr[VN] = ussr[RHO] * SM;
      ussl[VT] = ussl[RHO] * vss;
      ussr[VT] = ussr[RHO] * vss;
      ussl[VB] = ussl[RHO] * wss;
      ussr[VB] = ussr[RHO] * wss;
      const double nbt2 = sqrL * usR[BT] + sqrR * usL[BT] + sqrL * sqrR * (vsR - vsL) * sBx;
      const double nbb2 = sqrL * usR[BB] + sqrR * usL[BB] + sqrL * sqrR * (wsR - wsL) * sBx;
      uss[BT] = nbt2 / den;
      uss[BB] = nbb2 / den;
      ussl[BT] = ussr[BT] = uss[BT];
      ussl[BB] = ussr[BB] = uss[BB];
      ussl[BN] = ussr[BN] = Bx1;
      const double xL = (SM * Bx1 + vsL * usL[BT] + wsL * usL[BB]) -
                        (SM * Bx1 + vss * uss[BT] + wss * uss[BB]);
      const double xR = (SM * Bx1 + vsR * usR[BT] + wsR * usR[BB]) -
                        (SM * Bx1 + vss * uss[BT] + wss * uss[BB]);
      ussl[ENG] = usL[ENG] - sqrL * xL * sBx;
      ussr[ENG] = usR[ENG] + sqrR * xR * sBx;

      double a_ussl[NV] = {0}, a_ussr[NV] = {0};
      if (SM >= 0.0) {
#pragma unroll
        for (int q = 0; q < 5; q++) {
          const int nv = OUT[q];
          const double a = s.f[nv];
          a_fL[nv] += a;
          a_S1L += a * (ussl[nv] - usL[nv]);
          a_ussl[nv] += a * S1L;
          a_usL[nv] += a * (SL - S1L);
          a_SL += a * (usL[nv] - uL[nv]);
          a_uL[nv] -= a * SL;
        }
        a_ptL += s.press;
      } else {
#pragma unroll
        for (int q = 0; q < 5; q++) {
          const int nv = OUT[q];
          const double a = s.f[nv];
          a_fR[nv] += a;
          a_S1R += a * (ussr[nv] - usR[nv]);
          a_ussr[nv] += a * S1R;
          a_usR[nv] += a * (SR - S1R);
          a_SR += a * (usR[nv] - uR[nv]);
          a_uR[nv] -= a * SR;
        }
        a_ptR += s.press;
      }
      // reverse F18
      a_usL[ENG] += a_ussl[ENG];
      const double a_xL = -a_ussl[ENG] * sqrL * sBx;
      a_sqrL -= a_ussl[ENG] * xL * sBx;
      a_usR[ENG] += a_ussr[ENG];
      const double a_xR = a_ussr[ENG] * sqrR * sBx;
      a_sqrR += a_ussr[ENG] * xR * sBx;
      double a_vss = 0.0, a_wss = 0.0;
      double a_ubt = a_ussl[BT] + a_ussr[BT], a_ubb = a_ussl[BB] + a_ussr[BB];
      a_vsL += a_xL * usL[BT];
      a_usL[BT] += a_xL * vsL;
      a_wsL += a_xL * usL[BB];
      a_usL[BB] += a_xL * wsL;
      a_vsR += a_xR * usR[BT];
      a_usR[BT] += a_xR * vsR;
      a_wsR += a_xR * usR[BB];
      a_usR[BB] += a_xR * wsR;
      a_vss -= (a_xL + a_xR) * uss[BT];
      a_wss -= (a_xL + a_xR) * uss[BB];
      a_ubt -= (a_xL + a_xR) * vss;
      a_ubb -= (a_xL + a_xR) * wss;
      // reverse F17
      double a_den = 0.0;
      {
        const double q = a_ubt / den;
        a_sqrL += q * (usR[BT] + sqrR * (vsR - vsL) * sBx);
        a_sqrR += q * (usL[BT] + sqrL * (vsR - vsL) * sBx);
        a_usR[BT] += q * sqrL;
        a_usL[BT] += q * sqrR;
        a_vsR += q * sqrL * sqrR * sBx;
        a_vsL -= q * sqrL * sqrR * sBx;
        a_den -= q * uss[BT];
      }
      {
        const double q = a_ubb / den;
        a_sqrL += q * (usR[BB] + sqrR * (wsR - wsL) * sBx);
        a_sqrR += q * (usL[BB] + sqrL * (wsR - wsL) * sBx);
        a_usR[BB] += q * sqrL;
        a_usL[BB] += q * sqrR;
        a_wsR += q * sqrL * sqrR * sBx;
        a_wsL -= q * sqrL * sqrR * sBx;
        a_den -= q * uss[BB];
      }
      // reverse F16 (ussl[BN] = Bx1 is not an output, no adjoint)
      double a_urhoL = a_ussl[RHO], a_urhoR = a_ussr[RHO];
      a_urhoL += a_ussl[VN] * SM + a_ussl[VT] * vss + a_ussl[VB] * wss;
      a_urhoR += a_ussr[VN] * SM + a_ussr[VT] * vss + a_ussr[VB] * wss;
      a_SM += a_ussl[VN] * ussl[RHO] + a_ussr[VN] * ussr[RHO];
      a_vss += a_ussl[VT] * ussl[RHO] + a_ussr[VT] * ussr[RHO];
      a_wss += a_ussl[VB] * ussl[RHO] + a_ussr[VB] * ussr[RHO];
      a_usL[RHO] += a_urhoL;
      a_usR[RHO] += a_urhoR;
      // reverse F15
      {
        const double q = a_vss / den;
        a_sqrL += q * vsL;
        a_vsL += q * sqrL;
        a_sqrR += q * vsR;
        a_vsR += q * sqrR;
        a_usR[BT] += q * sBx;
        a_usL[BT] -= q * sBx;
        a_den -= q * vss;
      }
      {
        const double q = a_wss / den;
        a_sqrL += q * wsL;
        a_wsL += q * sqrL;
        a_sqrR += q * wsR;
        a_wsR += q * sqrR;
        a_usR[BB] += q * sBx;
        a_usL[BB] -= q * sBx;
        a_den -= q * wss;
      }
      a_sqrL += a_den;
      a_sqrR += a_den;
    }

    // reverse F14 (energy of the star states)
    double a_pts = 0.0, a_Bx1 = 0.0, a_duL = 0.0, a_duR = 0.0, a_dL = 0.0, a_dR = 0.0;
    {
      const double q = a_usL[ENG] / dL;
      a_dL -= q * usL[ENG];
      a_duL += q * uL[ENG];
      a_uL[ENG] += q * duL;
      a_ptL -= q * vL[VN];
      a_vL[VN] -= q * ptL;
      a_pts += q * SM;
      a_SM += q * pts;
      a_Bx += q * cL;
      const double a_c = q * Bx;
      a_vL[VN] += a_c * Bx1;
      a_Bx1 += a_c * vL[VN];
      a_vL[VT] += a_c * uL[BT];
      a_uL[BT] += a_c * vL[VT];
      a_vL[VB] += a_c * uL[BB];
      a_uL[BB] += a_c * vL[VB];
      a_SM -= a_c * Bx1;
      a_Bx1 -= a_c * SM;
      a_vsL -= a_c * usL[BT];
      a_usL[BT] -= a_c * vsL;
      a_wsL -= a_c * usL[BB];
      a_usL[BB] -= a_c * wsL;
    }
    {
      const double q = a_usR[ENG] / dR;
      a_dR -= q * usR[ENG];
      a_duR += q * uR[ENG];
      a_uR[ENG] += q * duR;
      a_ptR -= q * vR[VN];
      a_vR[VN] -= q * ptR;
      a_pts += q * SM;
      a_SM += q * pts;
      a_Bx += q * cR;
      const double a_c = q * Bx;
      a_vR[VN] += a_c * Bx1;
      a_Bx1 += a_c * vR[VN];
      a_vR[VT] += a_c * uR[BT];
      a_uR[BT] += a_c * vR[VT];
      a_vR[VB] += a_c * uR[BB];
      a_uR[BB] += a_c * vR[VB];
      a_SM -= a_c * Bx1;
      a_Bx1 -= a_c * SM;
      a_vsR -= a_c * usR[BT];
      a_usR[BT] -= a_c * vsR;
      a_wsR -= a_c * usR[BB];
      a_usR[BB] -= a_c * wsR;
    }
    // reverse F13
    a_usL[RHO] += a_usL[VN] * SM + a_usL[VT] * vsL + a_usL[VB] * wsL;
    a_usR[RHO] += a_usR[VN] * SM + a_usR[VT] * vsR + a_usR[VB] * wsR;
    a_SM += a_usL[VN] * usL[RHO] + a_usR[VN] * usR[RHO];
    a_vsL += a_usL[VT] * usL[RHO];
    a_vsR += a_usR[VT] * usR[RHO];
    a_wsL += a_usL[VB] * usL[RHO];
    a_wsR += a_usR[VB] * usR[RHO];
    // reverse F12
    double a_scrhL = 0.0, a_scrhR = 0.0;
    a_vL[VT] += a_vsL;
    a_scrhL -= a_vsL * (usL[BT] - uL[BT]);
    a_usL[BT] -= a_vsL * scrhL;
    a_uL[BT] += a_vsL * scrhL;
    a_vL[VB] += a_wsL;
    a_scrhL -= a_wsL * (usL[BB] - uL[BB]);
    a_usL[BB] -= a_wsL * scrhL;
    a_uL[BB] += a_wsL * scrhL;
    a_vR[VT] += a_vsR;
    a_scrhR -= a_vsR * (usR[BT] - uR[BT]);
    a_usR[BT] -= a_vsR * scrhR;
    a_uR[BT] += a_vsR * scrhR;
    a_vR[VB] += a_wsR;
    a_scrhR -= a_wsR * (usR[BB] - uR[BB]);
    a_usR[BB] -= a_wsR * scrhR;
    a_uR[BB] += a_wsR * scrhR;
    // reverse F11
    {
      a_Bx += a_scrhL / ddL;
      const double a_dd = -a_scrhL * scrhL / ddL;
      a_uL[RHO] += a_dd * duL;
      a_duL += a_dd * uL[RHO];
    }
    {
      a_Bx += a_scrhR / ddR;
      const double a_dd = -a_scrhR * scrhR / ddR;
      a_uR[RHO] += a_dd * duR;
      a_duR += a_dd * uR[RHO];
    }
    // reverse F10
    if (revert) {
      const double a_hn = a_usL[BN] + a_usR[BN];
      const double a_ht = a_usL[BT] + a_usR[BT];
      const double a_hb = a_usL[BB] + a_usR[BB];
      double a_scrh3 = 0.0;
      const int comps[3] = {BN, BT, BB};
      const double ah[3] = {a_hn, a_ht, a_hb};
      const double nh[3] = {nbn, nbt, nbb};
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const int c = comps[q];
        const double a = ah[q] * scrh3;
        a_SR += a * uR[c];
        a_uR[c] += a * SR;
        a_SL -= a * uL[c];
        a_uL[c] -= a * SL;
        a_fL[c] += a;
        a_fR[c] -= a;
        a_scrh3 += ah[q] * nh[q];
      }
      const double a_t3 = -a_scrh3 * scrh3 * scrh3;
      a_SR += a_t3;
      a_SL -= a_t3;
      a_SM += a_S1L + a_S1R;   // S1L = S1R = SM
    } else {
      a_Bx1 += a_usL[BN] + a_usR[BN];
      double a_sL = a_usL[BT] * uL[BT] + a_usL[BB] * uL[BB];
      double a_sR = a_usR[BT] * uR[BT] + a_usR[BB] * uR[BB];
      a_uL[BT] += a_usL[BT] * sL;
      a_uL[BB] += a_usL[BB] * sL;
      a_uR[BT] += a_usR[BT] * sR;
      a_uR[BB] += a_usR[BB] * sR;
      {
        const double a_n = a_sL / sLd, a_d = -a_sL * sL / sLd;
        a_uL[RHO] += a_n * duL * duL + a_d * duL * dL;
        a_duL += a_n * 2.0 * uL[RHO] * duL + a_d * uL[RHO] * dL;
        a_dL += a_d * uL[RHO] * duL;
        a_Bx -= (a_n + a_d) * 2.0 * Bx;
      }
      {
        const double a_n = a_sR / sRd, a_d = -a_sR * sR / sRd;
        a_uR[RHO] += a_n * duR * duR + a_d * duR * dR;
        a_duR += a_n * 2.0 * uR[RHO] * duR + a_d * uR[RHO] * dR;
        a_dR += a_d * uR[RHO] * duR;
        a_Bx -= (a_n + a_d) * 2.0 * Bx;
      }
      // reverse F9: S1L = SM - |Bx|/sqrL, S1R = SM + |Bx|/sqrR
      const double aB = fabs(Bx);
      a_SM += a_S1L + a_S1R;
      const double a_abs = -a_S1L / sqrL + a_S1R / sqrR;
      a_sqrL += a_S1L * aB / (sqrL * sqrL);
      a_sqrR -= a_S1R * aB / (sqrR * sqrR);
      a_Bx += a_abs * (Bx > 0.0 ? 1.0 : (Bx < 0.0 ? -1.0 : 0.0));
    }
    // reverse F8
    a_usL[RHO] += a_sqrL / (2.0 * sqrL);
    a_usR[RHO] += a_sqrR / (2.0 * sqrR);
    // reverse F7
    {
      const double q = a_usL[RHO] / dL;
      a_uL[RHO] += q * duL;
      a_duL += q * uL[RHO];
      a_dL -= q * usL[RHO];
    }
    {
      const double q = a_usR[RHO] / dR;
      a_uR[RHO] += q * duR;
      a_duR += q * uR[RHO];
      a_dR -= q * usR[RHO];
    }
    // dL = SL - SM, dR = SR - SM
    a_SL += a_dL;
    a_SM -= a_dL;
    a_SR += a_dR;
    a_SM -= a_dR;
    // reverse F6
    {
      const double a_n = a_pts * scrh2;
      double a_s2 = a_pts * num6;
      a_duR += a_n * (uR[RHO] * ptL + rr * duL * dvn);
      a_uR[RHO] += a_n * duR * ptL;
      a_ptL += a_n * duR * uR[RHO];
      a_duL += a_n * (-uL[RHO] * ptR + rr * duR * dvn);
      a_uL[RHO] -= a_n * duL * ptR;
      a_ptR -= a_n * duL * uL[RHO];
      a_vL[RHO] += a_n * vR[RHO] * duR * duL * dvn;
      a_vR[RHO] += a_n * vL[RHO] * duR * duL * dvn;
      a_vR[VN] += a_n * rr * duR * duL;
      a_vL[VN] -= a_n * rr * duR * duL;
      // reverse F5
      const double a_n5 = a_SM * scrh2;
      a_s2 += a_SM * num5;
      a_duR += a_n5 * uR[VN];
      a_uR[VN] += a_n5 * duR;
      a_duL -= a_n5 * uL[VN];
      a_uL[VN] -= a_n5 * duL;
      a_ptR -= a_n5;
      a_ptL += a_n5;
      // reverse F4
      const double a_den2 = -a_s2 * scrh2 * scrh2;
      a_duR += a_den2 * uR[RHO];
      a_uR[RHO] += a_den2 * duR;
      a_duL -= a_den2 * uL[RHO];
      a_uL[RHO] -= a_den2 * duL;
    }
    // reverse F3
    a_SL += a_duL;
    a_vL[VN] -= a_duL;
    a_SR += a_duR;
    a_vR[VN] -= a_duR;
    // reverse F2, F1
    a_Bx1 += a_Bx;
    {
      const double a_n = a_Bx1 * scrh1;
      const double a_s1 = a_Bx1 * num2;
      a_SR += a_n * vR[BN];
      a_vR[BN] += a_n * SR;
      a_SL -= a_n * vL[BN];
      a_vL[BN] -= a_n * SL;
      const double a_t1 = -a_s1 * scrh1 * scrh1;
      a_SR += a_t1;
      a_SL -= a_t1;
    }
  }
  // SL = min(slmin, srmin), SR = max(slmax, srmax)
  double a_slmin = 0.0, a_slmax = 0.0, a_srmin = 0.0, a_srmax = 0.0;
  if (selL) a_slmin = a_SL; else a_srmin = a_SL;
  if (selR) a_slmax = a_SR; else a_srmax = a_SR;
  fast_speeds_vjp<D>(vL, gamma, a_slmin, a_slmax, a_vL);
  fast_speeds_vjp<D>(vR, gamma, a_srmin, a_srmax, a_vR);
  normal_flux_vjp<D>(vL, uL, a_fL, a_ptL, a_vL, a_uL);
  normal_flux_vjp<D>(vR, uR, a_fR, a_ptR, a_vR, a_uR);
  prim2cons_vjp(vL, a_uL, a_vL, gmm1);
  prim2cons_vjp(vR, a_uR, a_vR, gmm1);
}

// UCT_HLL edge EMF (ct_emf_average.c:285-289) VJP: inputs in the order
// (a_xp, a_xm, a_yp, a_ym, eSW, eSE, eNE, eNW, bS, bN, bW, bE).
__device__ __forceinline__ void uct_hll_vjp(const double* x, double a_e, double* a_x)
{
  const double axp = x[0], axm = x[1], ayp = x[2], aym = x[3];
  const double eSW = x[4], eSE = x[5], eNE = x[6], eNW = x[7];
  const double bS = x[8], bN = x[9], bW = x[10], bE = x[11];
  const double sx = axp + axm, sy = ayp + aym;
  const double D = sx * sy;
  const double num = axp * ayp * eSW + axm * ayp * eSE + axm * aym * eNE + axp * aym * eNW;
  const double q = a_e / D;
  // e1 = num / D
  a_x[4] += q * axp * ayp;
  a_x[5] += q * axm * ayp;
  a_x[6] += q * axm * aym;
  a_x[7] += q * axp * aym;
  double g_axp = q * (ayp * eSW + aym * eNW);
  double g_axm = q * (ayp * eSE + aym * eNE);
  double g_ayp = q * (axp * eSW + axm * eSE);
  double g_aym = q * (axm * eNE + axp * eNW);
  const double e1 = num / D;
  g_axp -= q * e1 * sy;
  g_axm -= q * e1 * sy;
  g_ayp -= q * e1 * sx;
  g_aym -= q * e1 * sx;
  // e2 = ayp aym (bN - bS) / sy, subtracted
  {
    const double db = bN - bS;
    const double t = a_e / sy;
    a_x[9] -= t * ayp * aym;
    a_x[8] += t * ayp * aym;
    const double e2 = ayp * aym * db / sy;
    g_ayp -= t * aym * db - t * e2;
    g_aym -= t * ayp * db - t * e2;
  }
  // e3 = axp axm (bE - bW) / sx, added
  {
    const double db = bE - bW;
    const double t = a_e / sx;
    a_x[11] += t * axp * axm;
    a_x[10] -= t * axp * axm;
    const double e3 = axp * axm * db / sx;
    g_axp += t * axm * db - t * e3;
    g_axm += t * axp * db - t * e3;
  }
  a_x[0] += g_axp;
  a_x[1] += g_axm;
  a_x[2] += g_ayp;
  a_x[3] += g_aym;
}

// MC_LIM2 (ct_stag_slopes.c:92-106) VJP.
__device__ __forceinline__ void mc_lim2_vjp(double dp, double dm, double a_s, double& a_dp,
                                            double& a_dm)
{
  if (dp * dm < 0.0) return;
  const double dc = 0.5 * (dp + dm);
  const bool p_small = fabs(dp) < fabs(dm);
  const double sc = 2.0 * (p_small ? dp : dm);
  if (fabs(dc) < fabs(sc)) {
    a_dp += 0.5 * a_s;
    a_dm += 0.5 * a_s;
  } else if (p_small) {
    a_dp += 2.0 * a_s;
  } else {
    a_dm += 2.0 * a_s;
  }
}

#pragma clang fp contract(off)

}  // namespace pgpu
