// Device-side MHD building blocks for the unsplit stage on gfx950.
//
// Point functions used by the stage kernels: primitive/conservative maps,
// the MC-limited PLM slope, the ideal-MHD normal flux, the fast magnetosonic
// speed and the HLLD Riemann solver.  Every expression keeps PLUTO's operand
// order so that, compiled without FMA contraction, each value is the IEEE
// FP64 result the reference computes:
//   PrimToCons      MHD/mappers.c:25-86
//   ConsToPrim      MHD/mappers.c:88-254 (floors 1e-12, flag bit 64)
//   MC limiter      States/plm_coeffs.h:119-123, States/plm_states.c:233-240
//   Flux            MHD/fluxes.c:159-215
//   MaxSignalSpeed  MHD/eigenv.c:35-104, HLL_Speed MHD/hll_speed.c:76-107
//   HLLD_Solver     MHD/hlld.c:44-452
// The sweep direction is a template parameter: normal/tangent/binormal
// component indices follow SetVectorIndices (set_indexes.c:49-123), i.e.
// x -> (x,y,z), y -> (y,x,z), z -> (z,x,y).
//
// The Riemann-solver arithmetic is templated on the real type R so the same
// code serves the FP64 path and the FP32-flux path; conversions to and from
// the FP64 state happen at the call sites.

#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgpu {

constexpr int NV = 8;
enum : int { RHO = 0, VX1 = 1, VX2 = 2, VX3 = 3, BX1 = 4, BX2 = 5, BX3 = 6, PRS = 7 };
constexpr int ENG = PRS;
constexpr unsigned char FLAG_CONS2PRIM_FAIL = 64;

template <int D> struct Dir;
template <> struct Dir<0> { static constexpr int n = 0, t = 1, b = 2; };
template <> struct Dir<1> { static constexpr int n = 1, t = 0, b = 2; };
template <> struct Dir<2> { static constexpr int n = 2, t = 0, b = 1; };

// Division and square root of the Riemann-solver arithmetic.  FP64 keeps the
// IEEE operations (bit-exact with the reference); the FP32-flux path uses the
// hardware reciprocal and square root (about 1 ulp), which its error budget
// (relative 1e-6 against FP64) absorbs many times over.
//
// A translation unit that defines PGPU_FAST_F64 before including this header
// (fused_fast.hip, the tolerance-mode FP64 flux kernels, with their own
// kernel symbols) gets FP64 division and square root by Newton iteration
// instead: the reciprocal estimate v_rcp_f64 refined twice (e <- 1 - b r,
// r <- r + r e) and the quotient corrected once (q <- q + r (a - b q)); the
// reciprocal square root estimate v_rsq_f64 refined by one coupled
// Goldschmidt step for sqrt and 1/(2 sqrt), then one Newton correction of
// the root.  Both are within an ulp or two of the IEEE results for the
// normal, finite, positive operands the solvers divide by and take roots of.
template <typename R> __device__ __forceinline__ R div_(R a, R b);
#ifdef PGPU_FAST_F64
template <> __device__ __forceinline__ double div_(double a, double b)
{
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(e, r, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(e, r, r);
  double q = a * r;
  const double rem = __builtin_fma(-b, q, a);
  return __builtin_fma(rem, r, q);
}
#else
template <> __device__ __forceinline__ double div_(double a, double b) { return a / b; }
#endif
template <> __device__ __forceinline__ float div_(float a, float b)
{
  return a * __builtin_amdgcn_rcpf(b);
}
template <typename R> __device__ __forceinline__ R sqrt_(R x);
#ifdef PGPU_FAST_F64
template <> __device__ __forceinline__ double sqrt_(double x)
{
  const double y = __builtin_amdgcn_rsq(x);
  double s = x * y, h = 0.5 * y;
  const double e = __builtin_fma(-s, h, 0.5);
  s = __builtin_fma(s, e, s);
  h = __builtin_fma(h, e, h);
  const double d = __builtin_fma(-s, s, x);
  s = __builtin_fma(d, h, s);
  return x > 0.0 ? s : x;   // sqrt(0) = 0 (the estimate is infinite there)
}
#else
template <> __device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
#endif
template <> __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }

// Branch recorder of the step's piecewise operations (limiter cases, min
// and max, Riemann-solver regions): every comparison whose outcome picks
// which expression a value follows folds into a running 64-bit hash.  Two
// runs whose hashes agree took the same branches everywhere, i.e. their
// results lie on one smooth piece of the step.  With a tape mode the
// outcomes also go, one bit each in the order they occur, into `word`
// (mode 1), or are taken from it instead of the comparison (mode 2): the
// replayed computation is the smooth piece of the recorded run, evaluated
// at the current inputs.  The functions below take an optional recorder
// that defaults to null; with null (every production kernel) the recording
// code compiles away.
struct BranchRec {
  unsigned long long h = 0;     // running hash of the outcomes
  unsigned long long word = 0;  // tape: outcome bits
  int pos = 0;                  // tape: next bit
  int mode = 0;                 // 0 hash only, 1 record into word, 2 replay from word
};

__device__ __forceinline__ bool rb(BranchRec* r, bool c)
{
  if (r) {
    if (r->mode == 2) c = ((r->word >> r->pos) & 1ull) != 0;
    else if (r->mode == 1 && c) r->word |= 1ull << r->pos;
    r->pos++;
    r->h = (r->h ^ (c ? 0x9E3779B97F4A7C15ull : 0xC2B2AE3D27D4EB4Full)) * 0x100000001B3ull;
  }
  return c;
}

template <typename R>
__device__ __forceinline__ R abs_min(R a, R b, BranchRec* rec = nullptr)
{
  return rb(rec, fabs(a) < fabs(b)) ? a : b;
}

template <typename R>
__device__ __forceinline__ R max_ge(R a, R b, BranchRec* rec = nullptr)
{
  return rb(rec, a >= b) ? a : b;
}

template <typename R>
__device__ __forceinline__ R min_le(R a, R b, BranchRec* rec = nullptr)
{
  return rb(rec, a <= b) ? a : b;
}

// MC-limited undivided slope of a cell from its two neighbours.
template <typename R>
__device__ __forceinline__ R mc_slope(R vm1, R v0, R vp1, BranchRec* rec = nullptr)
{
  R dvp = vp1 - v0;
  R dvm = v0 - vm1;
  if (rb(rec, dvp * dvm > R(0))) {
    R qc = R(0.5) * (dvm + dvp);
    R sc = R(2.0) * abs_min(dvp, dvm, rec);
    return abs_min(qc, sc, rec);
  }
  return R(0);
}

template <typename R>
__device__ __forceinline__ void prim2cons(const R* v, R* u, R gmm1)
{
  u[RHO] = v[RHO];
  u[VX1] = v[RHO] * v[VX1];
  u[VX2] = v[RHO] * v[VX2];
  u[VX3] = v[RHO] * v[VX3];
  u[BX1] = v[BX1];
  u[BX2] = v[BX2];
  u[BX3] = v[BX3];
  R kinb2 = v[VX1] * v[VX1] + v[VX2] * v[VX2] + v[VX3] * v[VX3];
  kinb2 = v[RHO] * kinb2 + v[BX1] * v[BX1] + v[BX2] * v[BX2] + v[BX3] * v[BX3];
  kinb2 *= R(0.5);
  u[ENG] = kinb2 + div_(v[PRS], gmm1);
}

// Total energy only (the other conservative components are products).
__device__ __forceinline__ double total_energy(const double* v, double gmm1)
{
  double kinb2 = v[VX1] * v[VX1] + v[VX2] * v[VX2] + v[VX3] * v[VX3];
  kinb2 = v[RHO] * kinb2 + v[BX1] * v[BX1] + v[BX2] * v[BX2] + v[BX3] * v[BX3];
  kinb2 *= 0.5;
  return kinb2 + v[PRS] / gmm1;
}

// PLUTO's CheckNaN (tools.c:20-38) tests every variable of the stage input
// for v != v; the primitive state a stage writes is the next stage's input.
__device__ __forceinline__ bool has_nan(const double* v)
{
  bool bad = false;
#pragma unroll
  for (int nv = 0; nv < NV; nv++) bad |= (v[nv] != v[nv]);
  return bad;
}

// ConsToPrim with the reference's floors; u is modified in place by floors.
__device__ __forceinline__ unsigned char cons2prim(double* u, double* v, double gmm1)
{
  const double small_density = 1.e-12, small_pressure = 1.e-12;
  unsigned char flag = 0;
  double m2 = u[VX1] * u[VX1] + u[VX2] * u[VX2] + u[VX3] * u[VX3];
  double b2 = u[BX1] * u[BX1] + u[BX2] * u[BX2] + u[BX3] * u[BX3];
  if (u[RHO] < 0.0) {
    u[RHO] = small_density;
    flag |= FLAG_CONS2PRIM_FAIL;
  }
  v[RHO] = u[RHO];
  double tau = 1.0 / u[RHO];
  v[VX1] = u[VX1] * tau;
  v[VX2] = u[VX2] * tau;
  v[VX3] = u[VX3] * tau;
  v[BX1] = u[BX1];
  v[BX2] = u[BX2];
  v[BX3] = u[BX3];
  double kinb2 = 0.5 * (m2 * tau + b2);
  if (u[ENG] < 0.0) {
    u[ENG] = small_pressure / gmm1 + kinb2;
    flag |= FLAG_CONS2PRIM_FAIL;
  }
  v[PRS] = gmm1 * (u[ENG] - kinb2);
  if (v[PRS] < 0.0) {
    v[PRS] = small_pressure;
    u[ENG] = v[PRS] / gmm1 + kinb2;
    flag |= FLAG_CONS2PRIM_FAIL;
  }
  return flag;
}

// Normal flux of one state.  FULL also produces the induction components,
// which the stage kernels never consume (cell-centred B is rebuilt from the
// staggered faces and the UCT_HLL EMF does not use them).
template <int D, bool FULL, typename R>
__device__ __forceinline__ void normal_flux(const R* v, const R* u, R* fx, R& ptot)
{
  constexpr int VN = VX1 + Dir<D>::n, VT = VX1 + Dir<D>::t, VB = VX1 + Dir<D>::b;
  constexpr int BN = BX1 + Dir<D>::n, BT = BX1 + Dir<D>::t, BB = BX1 + Dir<D>::b;
  R bmag2 = v[BX1] * v[BX1] + v[BX2] * v[BX2] + v[BX3] * v[BX3];
  ptot = v[PRS] + R(0.5) * bmag2;
  R vb = v[VX1] * v[BX1] + v[VX2] * v[BX2] + v[VX3] * v[BX3];
  fx[RHO] = u[VN];
  fx[VX1] = v[VN] * u[VX1] - v[BN] * v[BX1];
  fx[VX2] = v[VN] * u[VX2] - v[BN] * v[BX2];
  fx[VX3] = v[VN] * u[VX3] - v[BN] * v[BX3];
  if (FULL) {
    fx[BN] = R(0);
    fx[BT] = v[VN] * v[BT] - v[BN] * v[VT];
    fx[BB] = v[VN] * v[BB] - v[BN] * v[VB];
  }
  fx[ENG] = (u[ENG] + ptot) * v[VN] - v[BN] * vb;
}

template <int D, typename R>
__device__ __forceinline__ void fast_speeds(const R* v, R gamma, R& cmin, R& cmax)
{
  constexpr int VN = VX1 + Dir<D>::n;
  constexpr int BN = BX1 + Dir<D>::n, BT = BX1 + Dir<D>::t, BB = BX1 + Dir<D>::b;
  R gpr = gamma * v[PRS];
  R b1 = v[BN], b2 = v[BT], b3 = v[BB];
  R btmag2 = b2 * b2 + b3 * b3;
  R bmag2 = b1 * b1 + btmag2;
  R cf = gpr - bmag2;
  cf = gpr + bmag2 + sqrt_(cf * cf + R(4.0) * gpr * btmag2);
  cf = sqrt_(div_(R(0.5) * cf, v[RHO]));
  cmin = v[VN] - cf;
  cmax = v[VN] + cf;
}

template <typename R>
struct RiemannOut {
  R f[NV];     // flux; f[BX*] valid only for FULL
  R press;     // total pressure at the interface (separate gradient term)
  R SL, SR;    // Davis estimates of the outer waves
  R cmax;      // max(|SL|, |SR|)
  R mach;      // (|vL_n| + |vR_n|) / (cs_L + cs_R)
};

// HLLD of Miyoshi & Kusano at one interface; vL/vR are primitive states
// whose normal field already holds the staggered face value.
template <int D, bool FULL, typename R>
__device__ __forceinline__ void hlld(const R* vL, const R* vR, R gamma, R gmm1,
                                     RiemannOut<R>& o, BranchRec* rec = nullptr)
{
  constexpr int VN = VX1 + Dir<D>::n, VT = VX1 + Dir<D>::t, VB = VX1 + Dir<D>::b;
  constexpr int BN = BX1 + Dir<D>::n, BT = BX1 + Dir<D>::t, BB = BX1 + Dir<D>::b;
  R uL[NV], uR[NV], fL[NV], fR[NV], ptL, ptR;

  prim2cons(vL, uL, gmm1);
  prim2cons(vR, uR, gmm1);
  R a2L = div_(gamma * vL[PRS], vL[RHO]);
  R a2R = div_(gamma * vR[PRS], vR[RHO]);
  normal_flux<D, true>(vL, uL, fL, ptL);
  normal_flux<D, true>(vR, uR, fR, ptR);

  R slmin, slmax, srmin, srmax;
  fast_speeds<D>(vL, gamma, slmin, slmax);
  fast_speeds<D>(vR, gamma, srmin, srmax);
  const R SL = min_le(slmin, srmin, rec);
  const R SR = max_ge(slmax, srmax, rec);
  {
    R m = fabs(vL[VN]) + fabs(vR[VN]);
    m = div_(m, sqrt_(a2L) + sqrt_(a2R));
    o.mach = m;
  }
  o.SL = SL;
  o.SR = SR;
  o.cmax = max_ge(fabs(SL), fabs(SR));

  if (rb(rec, SL >= R(0))) {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) o.f[nv] = fL[nv];
    o.press = ptL;
    return;
  }
  if (rb(rec, SR <= R(0))) {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) o.f[nv] = fR[nv];
    o.press = ptR;
    return;
  }

  R usL[NV], usR[NV];
  R scrh = div_(R(1), SR - SL);
  const R Bx1 = (SR * vR[BN] - SL * vL[BN]) * scrh;
  const R Bx = Bx1;
  const R sBx = (rb(rec, Bx > R(0)) ? R(1) : R(-1));
  // |Bx|; a replayed run stays on the recorded side of its kink: sBx * Bx
  const R aBx = (rec && rec->mode == 2) ? sBx * Bx : fabs(Bx);

  const R duL = SL - vL[VN];
  const R duR = SR - vR[VN];

  scrh = div_(R(1), duR * uR[RHO] - duL * uL[RHO]);
  const R SM = (duR * uR[VN] - duL * uL[VN] - ptR + ptL) * scrh;

  R pts = duR * uR[RHO] * ptL - duL * uL[RHO] * ptR +
          vL[RHO] * vR[RHO] * duR * duL * (vR[VN] - vL[VN]);
  pts *= scrh;

  usL[RHO] = div_(uL[RHO] * duL, SL - SM);
  usR[RHO] = div_(uR[RHO] * duR, SR - SM);

  const R sqrL = sqrt_(usL[RHO]);
  const R sqrR = sqrt_(usR[RHO]);

  R S1L = SM - div_(aBx, sqrL);
  R S1R = SM + div_(aBx, sqrR);

  bool revert = false;
  if (rb(rec, (S1L - SL) < R(1.e-4) * (SM - SL))) revert = true;
  if (rb(rec, (S1R - SR) > -R(1.e-4) * (SR - SM))) revert = true;

  if (revert) {
    scrh = div_(R(1), SR - SL);
    R hbn = SR * uR[BN] - SL * uL[BN] + fL[BN] - fR[BN];
    R hbt = SR * uR[BT] - SL * uL[BT] + fL[BT] - fR[BT];
    R hbb = SR * uR[BB] - SL * uL[BB] + fL[BB] - fR[BB];
    hbn *= scrh;
    hbt *= scrh;
    hbb *= scrh;
    usL[BN] = usR[BN] = hbn;
    usL[BT] = usR[BT] = hbt;
    usL[BB] = usR[BB] = hbb;
    S1L = S1R = SM;
  } else {
    R sL = div_(uL[RHO] * duL * duL - Bx * Bx, uL[RHO] * duL * (SL - SM) - Bx * Bx);
    R sR = div_(uR[RHO] * duR * duR - Bx * Bx, uR[RHO] * duR * (SR - SM) - Bx * Bx);
    usL[BN] = Bx1;
    usL[BT] = uL[BT] * sL;
    usL[BB] = uL[BB] * sL;
    usR[BN] = Bx1;
    usR[BT] = uR[BT] * sR;
    usR[BB] = uR[BB] * sR;
  }

  R scrhL = div_(Bx, uL[RHO] * duL);
  R scrhR = div_(Bx, uR[RHO] * duR);

  const R vsL = vL[VT] - scrhL * (usL[BT] - uL[BT]);
  const R vsR = vR[VT] - scrhR * (usR[BT] - uR[BT]);
  const R wsL = vL[VB] - scrhL * (usL[BB] - uL[BB]);
  const R wsR = vR[VB] - scrhR * (usR[BB] - uR[BB]);

  usL[VN] = usL[RHO] * SM;
  usR[VN] = usR[RHO] * SM;
  usL[VT] = usL[RHO] * vsL;
  usR[VT] = usR[RHO] * vsR;
  usL[VB] = usL[RHO] * wsL;
  usR[VB] = usR[RHO] * wsR;

  scrhL = vL[VN] * Bx1 + vL[VT] * uL[BT] + vL[VB] * uL[BB];
  scrhL -= SM * Bx1 + vsL * usL[BT] + wsL * usL[BB];
  usL[ENG] = duL * uL[ENG] - ptL * vL[VN] + pts * SM + Bx * scrhL;
  usL[ENG] = div_(usL[ENG], SL - SM);

  scrhR = vR[VN] * Bx1 + vR[VT] * uR[BT] + vR[VB] * uR[BB];
  scrhR -= SM * Bx1 + vsR * usR[BT] + wsR * usR[BB];
  usR[ENG] = duR * uR[ENG] - ptR * vR[VN] + pts * SM + Bx * scrhR;
  usR[ENG] = div_(usR[ENG], SR - SM);

  if (rb(rec, S1L >= R(0))) {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      o.f[nv] = fL[nv] + SL * (usL[nv] - uL[nv]);
    }
    o.press = ptL;
    return;
  }
  if (rb(rec, S1R <= R(0))) {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      o.f[nv] = fR[nv] + SR * (usR[nv] - uR[nv]);
    }
    o.press = ptR;
    return;
  }

  R ussl[NV], ussr[NV];
  ussl[RHO] = usL[RHO];
  ussr[RHO] = usR[RHO];

  R vss = sqrL * vsL + sqrR * vsR + (usR[BT] - usL[BT]) * sBx;
  vss = div_(vss, sqrL + sqrR);
  R wss = sqrL * wsL + sqrR * wsR + (usR[BB] - usL[BB]) * sBx;
  wss = div_(wss, sqrL + sqrR);

  ussl[VN] = ussl[RHO] * SM;
  ussr[VN] = ussr[RHO] * SM;
  ussl[VT] = ussl[RHO] * vss;
  ussr[VT] = ussr[RHO] * vss;
  ussl[VB] = ussl[RHO] * wss;
  ussr[VB] = ussr[RHO] * wss;

  ussl[BN] = ussr[BN] = Bx1;
  ussl[BT] = sqrL * usR[BT] + sqrR * usL[BT] + sqrL * sqrR * (vsR - vsL) * sBx;
  ussl[BT] = div_(ussl[BT], sqrL + sqrR);
  ussr[BT] = ussl[BT];
  ussl[BB] = sqrL * usR[BB] + sqrR * usL[BB] + sqrL * sqrR * (wsR - wsL) * sBx;
  ussl[BB] = div_(ussl[BB], sqrL + sqrR);
  ussr[BB] = ussl[BB];

  scrhL = SM * Bx1 + vsL * usL[BT] + wsL * usL[BB];
  scrhL -= SM * Bx1 + vss * ussl[BT] + wss * ussl[BB];
  scrhR = SM * Bx1 + vsR * usR[BT] + wsR * usR[BB];
  scrhR -= SM * Bx1 + vss * ussr[BT] + wss * ussr[BB];

  ussl[ENG] = usL[ENG] - sqrL * scrhL * sBx;
  ussr[ENG] = usR[ENG] + sqrR * scrhR * sBx;

  if (rb(rec, SM >= R(0))) {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      o.f[nv] = fL[nv] + S1L * (ussl[nv] - usL[nv]) + SL * (usL[nv] - uL[nv]);
    }
    o.press = ptL;
  } else {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      o.f[nv] = fR[nv] + S1R * (ussr[nv] - usR[nv]) + SR * (usR[nv] - uR[nv]);
    }
    o.press = ptR;
  }
}

// HLL (single intermediate state, Davis wave-speed estimates) at one
// interface (MHD/hll.c:30-169): the same Flux, SoundSpeed2 and HLL_Speed
// (with its Mach estimate) as hlld; the outer states are taken for SL > 0
// and SR < 0 strictly (hll.c:117,123), the intermediate flux is
// (SL SR (uR - uL) + SR fL - SL fR) / (SR - SL) in that operation order.
template <int D, bool FULL, typename R>
__device__ __forceinline__ void hll(const R* vL, const R* vR, R gamma, R gmm1, RiemannOut<R>& o)
{
  constexpr int VN = VX1 + Dir<D>::n;
  R uL[NV], uR[NV], fL[NV], fR[NV], ptL, ptR;
  prim2cons(vL, uL, gmm1);
  prim2cons(vR, uR, gmm1);
  R a2L = div_(gamma * vL[PRS], vL[RHO]);
  R a2R = div_(gamma * vR[PRS], vR[RHO]);
  normal_flux<D, true>(vL, uL, fL, ptL);
  normal_flux<D, true>(vR, uR, fR, ptR);
  R slmin, slmax, srmin, srmax;
  fast_speeds<D>(vL, gamma, slmin, slmax);
  fast_speeds<D>(vR, gamma, srmin, srmax);
  const R SL = min_le(slmin, srmin);
  const R SR = max_ge(slmax, srmax);
  {
    R m = fabs(vL[VN]) + fabs(vR[VN]);
    m = div_(m, sqrt_(a2L) + sqrt_(a2R));
    o.mach = m;
  }
  o.SL = SL;
  o.SR = SR;
  o.cmax = max_ge(fabs(SL), fabs(SR));
  if (SL > R(0)) {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      o.f[nv] = fL[nv];
    }
    o.press = ptL;
  } else if (SR < R(0)) {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      o.f[nv] = fR[nv];
    }
    o.press = ptR;
  } else {
    const R scrh = div_(R(1), SR - SL);
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      R f = SL * SR * (uR[nv] - uL[nv]) + SR * fL[nv] - SL * fR[nv];
      f *= scrh;
      o.f[nv] = f;
    }
    o.press = (SR * ptL - SL * ptR) * scrh;
  }
}

// HLLC for MHD after Li (2005) as PLUTO modifies it (MHD/hllc.c:32-230):
// the same Flux, SoundSpeed2 and HLL_Speed as hlld; outer states for
// SL >= 0 and SR <= 0; otherwise the HLL average state and flux give the
// contact speed vxs, the total pressure ps and the star field, and the
// star state on the side of the contact gives the flux.  The magnetic
// energy and v.B sums run over x, y, z in that order (EXPAND).
template <int D, bool FULL, typename R>
__device__ __forceinline__ void hllc(const R* vL, const R* vR, R gamma, R gmm1, RiemannOut<R>& o)
{
  constexpr int VN = VX1 + Dir<D>::n, VT = VX1 + Dir<D>::t, VB = VX1 + Dir<D>::b;
  constexpr int BN = BX1 + Dir<D>::n, BT = BX1 + Dir<D>::t, BB = BX1 + Dir<D>::b;
  R uL[NV], uR[NV], fL[NV], fR[NV], ptL, ptR;
  prim2cons(vL, uL, gmm1);
  prim2cons(vR, uR, gmm1);
  R a2L = div_(gamma * vL[PRS], vL[RHO]);
  R a2R = div_(gamma * vR[PRS], vR[RHO]);
  normal_flux<D, true>(vL, uL, fL, ptL);
  normal_flux<D, true>(vR, uR, fR, ptR);
  R slmin, slmax, srmin, srmax;
  fast_speeds<D>(vL, gamma, slmin, slmax);
  fast_speeds<D>(vR, gamma, srmin, srmax);
  const R SL = min_le(slmin, srmin);
  const R SR = max_ge(slmax, srmax);
  {
    R m = fabs(vL[VN]) + fabs(vR[VN]);
    m = div_(m, sqrt_(a2L) + sqrt_(a2R));
    o.mach = m;
  }
  o.SL = SL;
  o.SR = SR;
  o.cmax = max_ge(fabs(SL), fabs(SR));
  if (SL >= R(0)) {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      o.f[nv] = fL[nv];
    }
    o.press = ptL;
    return;
  }
  if (SR <= R(0)) {
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      o.f[nv] = fR[nv];
    }
    o.press = ptR;
    return;
  }
  const R scrh = div_(R(1), SR - SL);
  R Uh[NV], Fh[NV];
#pragma unroll
  for (int nv = 0; nv < NV; nv++) {
    R a = SR * uR[nv] - SL * uL[nv] + fL[nv] - fR[nv];
    a *= scrh;
    Uh[nv] = a;
    R f = SL * SR * (uR[nv] - uL[nv]) + SR * fL[nv] - SL * fR[nv];
    f *= scrh;
    Fh[nv] = f;
  }
  Uh[VN] += (ptL - ptR) * scrh;
  Fh[VN] += (SR * ptL - SL * ptR) * scrh;

  R pl = vL[BX1] * vL[BX1] + vL[BX2] * vL[BX2] + vL[BX3] * vL[BX3];
  R pr = vR[BX1] * vR[BX1] + vR[BX2] * vR[BX2] + vR[BX3] * vR[BX3];
  pl = vL[PRS] + R(0.5) * pl;
  pr = vR[PRS] + R(0.5) * pr;
  const R vBl = vL[VX1] * vL[BX1] + vL[VX2] * vL[BX2] + vL[VX3] * vL[BX3];
  const R vBr = vR[VX1] * vR[BX1] + vR[VX2] * vR[BX2] + vR[VX3] * vR[BX3];
  const R vxl = vL[VN], vxr = vR[VN];
  const R Bxs = Uh[BN], Bys = Uh[BT], Bzs = Uh[BB];
  const R vxs = div_(Uh[VN], Uh[RHO]);
  const R ps = Fh[VN] + Bxs * Bxs - Fh[RHO] * vxs;
  R vBs = Uh[BX1] * Uh[VX1] + Uh[BX2] * Uh[VX2] + Uh[BX3] * Uh[VX3];
  vBs = div_(vBs, Uh[RHO]);

  if (vxs >= R(0)) {
    R us[NV];
    us[RHO] = div_(uL[RHO] * (SL - vxl), SL - vxs);
    us[ENG] = div_(uL[ENG] * (SL - vxl) + ps * vxs - pl * vxl - Bxs * vBs + vL[BN] * vBl, SL - vxs);
    us[VN] = us[RHO] * vxs;
    us[VT] = div_(uL[VT] * (SL - vxl) - (Bxs * Bys - vL[BN] * vL[BT]), SL - vxs);
    us[VB] = div_(uL[VB] * (SL - vxl) - (Bxs * Bzs - vL[BN] * vL[BB]), SL - vxs);
    us[BN] = Bxs;
    us[BT] = Bys;
    us[BB] = Bzs;
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      o.f[nv] = fL[nv] + SL * (us[nv] - uL[nv]);
    }
    o.press = ptL;
  } else {
    R us[NV];
    us[RHO] = div_(uR[RHO] * (SR - vxr), SR - vxs);
    us[ENG] = div_(uR[ENG] * (SR - vxr) + ps * vxs - pr * vxr - Bxs * vBs + vR[BN] * vBr, SR - vxs);
    us[VN] = us[RHO] * vxs;
    us[VT] = div_(uR[VT] * (SR - vxr) - (Bxs * Bys - vR[BN] * vR[BT]), SR - vxs);
    us[VB] = div_(uR[VB] * (SR - vxr) - (Bxs * Bzs - vR[BN] * vR[BB]), SR - vxs);
    us[BN] = Bxs;
    us[BT] = Bys;
    us[BB] = Bzs;
#pragma unroll
    for (int nv = 0; nv < NV; nv++) {
      if (!FULL && nv >= BX1 && nv <= BX3) continue;
      o.f[nv] = fR[nv] + SR * (us[nv] - uR[nv]);
    }
    o.press = ptR;
  }
}

// Riemann solvers behind PLUTO's Riemann_Solver pointer (set_solver.c)
// Lax-Friedrichs (Rusanov) flux of PLUTO's LF_Solver, "tvdlf"
// (MHD/tvdlf.c:37-163): c = max |v_n -+ c_f| of the arithmetic mean state
// (MaxSignalSpeed), F = (F_L + F_R - c (U_R - U_L)) / 2, SL = -c, SR = c;
// the Mach estimate is |v_n| / c_s of the mean state.
template <int D, bool FULL, typename R>
__device__ __forceinline__ void tvdlf(const R* vL, const R* vR, R gamma, R gmm1, RiemannOut<R>& o)
{
  constexpr int VN = VX1 + Dir<D>::n;
  R uL[NV], uR[NV], fL[NV], fR[NV], ptL, ptR, vRL[NV];
  prim2cons(vL, uL, gmm1);
  prim2cons(vR, uR, gmm1);
  normal_flux<D, true>(vL, uL, fL, ptL);
  normal_flux<D, true>(vR, uR, fR, ptR);
#pragma unroll
  for (int nv = 0; nv < NV; nv++) vRL[nv] = R(0.5) * (vL[nv] + vR[nv]);
  o.mach = div_(fabs(vRL[VN]), sqrt_(div_(gamma * vRL[PRS], vRL[RHO])));
  R cmin, cmax;
  fast_speeds<D>(vRL, gamma, cmin, cmax);
  const R c = max_ge(fabs(cmin), fabs(cmax));
  o.SL = -c;
  o.SR = c;
  o.cmax = c;
#pragma unroll
  for (int nv = 0; nv < NV; nv++) {
    if (!FULL && nv >= BX1 && nv <= BX3) continue;
    o.f[nv] = R(0.5) * (fL[nv] + fR[nv] - c * (uR[nv] - uL[nv]));
  }
  o.press = R(0.5) * (ptL + ptR);
}

// Roe solver of PLUTO's Roe_Solver, "roe" (MHD/roe.c:53-700; Cargo & Gallice
// 1997): Roe averages, the conservative eigenvectors Rc and wave strengths
// eta of roe.c:383-604 in the wave order of mod_defs.h:54-68 (fast-, fast+,
// entropy, div.B, slow-, slow+, Alfven-, Alfven+), the entropy fix of
// roe.c:616-631 and F = (F_L + F_R - sum_k |lambda_k| eta_k Rc_k) / 2 with the
// sum in wave order; SL, SR = u -+ c_f; cmax = |u| + c_f; Mach |u / a|.
// HLL_HYBRIDIZATION is NO there (roe.c:49).  two_comp: the reference's
// COMPONENTS 2 build used for 2D boxes (beta_y = sign(By), or 1 when the
// transverse field vanishes, and no Alfven waves).  A negative a^2 stops the
// reference (roe.c:301-308); here it propagates as NaN.
template <int D, bool FULL, typename R>
__device__ __forceinline__ void roe(const R* vL, const R* vR, R gamma, R gmm1, RiemannOut<R>& o,
                                    bool two_comp)
{
  constexpr int KFM = 0, KFP = 1, KEN = 2, KDB = 3, KSM = 4, KSP = 5, KAM = 6, KAP = 7;
  constexpr int VXn = VX1 + Dir<D>::n, VXt = VX1 + Dir<D>::t, VXb = VX1 + Dir<D>::b;
  constexpr int MXn = VXn, MXt = VXt, MXb = VXb;
  constexpr int BXn = BX1 + Dir<D>::n, BXt = BX1 + Dir<D>::t, BXb = BX1 + Dir<D>::b;
  const R delta = R(1.e-6), g1 = gmm1, sqrt_1_2 = R(0.70710678118654752440);
  R uL[NV], uR[NV], fL[NV], fR[NV], pL, pR;
  prim2cons(vL, uL, gmm1);
  prim2cons(vR, uR, gmm1);
  normal_flux<D, true>(vL, uL, fL, pL);
  normal_flux<D, true>(vR, uR, fR, pR);
  R dV[NV], dU[NV], Rc[NV][NV], eta[NV], lambda[NV], alambda[NV];
#pragma unroll
  for (int nv = 0; nv < NV; nv++) {
    dV[nv] = vR[nv] - vL[nv];
    dU[nv] = uR[nv] - uL[nv];
#pragma unroll
    for (int k = 0; k < NV; k++) Rc[nv][k] = R(0);
  }
  const R sqr_rho_L = sqrt_(vL[RHO]);
  const R sqr_rho_R = sqrt_(vR[RHO]);
  const R sl = div_(sqr_rho_L, sqr_rho_L + sqr_rho_R);
  const R sr = div_(sqr_rho_R, sqr_rho_L + sqr_rho_R);
  const R rho = sr * vL[RHO] + sl * vR[RHO];
  const R sqrt_rho = sqrt_(rho);
  const R u = sl * vL[VXn] + sr * vR[VXn];
  const R v = sl * vL[VXt] + sr * vR[VXt];
  const R w = sl * vL[VXb] + sr * vR[VXb];
  const R Bx = sr * vL[BXn] + sl * vR[BXn];
  const R By = sr * vL[BXt] + sl * vR[BXt];
  const R Bz = sr * vL[BXb] + sl * vR[BXb];
  const R sBx = (Bx >= R(0) ? R(1) : R(-1));
  const R bx = div_(Bx, sqrt_rho);
  const R by = div_(By, sqrt_rho);
  const R bz = div_(Bz, sqrt_rho);
  const R bt2 = R(0) + by * by + bz * bz;
  const R b2 = bx * bx + bt2;
  const R Btmag = sqrt_(bt2 * rho);
  R X = dV[BXn] * dV[BXn] + dV[BXt] * dV[BXt] + dV[BXb] * dV[BXb];
  X = div_(X, (sqr_rho_L + sqr_rho_R) * (sqr_rho_L + sqr_rho_R) * R(2.0));
  const R vdm = u * dU[MXn] + v * dU[MXt] + w * dU[MXb];
  const R BdB = Bx * dU[BXn] + By * dU[BXt] + Bz * dU[BXb];
  const R vel2 = u * u + v * v + w * w;
  dV[PRS] = g1 * ((R(0.5) * vel2 - X) * dV[RHO] - vdm + dU[ENG] - BdB);
  const R HL = div_(uL[ENG] + pL, vL[RHO]);
  const R HR = div_(uR[ENG] + pR, vR[RHO]);
  const R H = sl * HL + sr * HR;
  const R Hgas = H - b2;
  const R a2 = (R(2.0) - gamma) * X + g1 * (Hgas - R(0.5) * vel2);

  R scrh = a2 - b2;
  const R ca2 = bx * bx;
  scrh = scrh * scrh + R(4.0) * bt2 * a2;
  scrh = sqrt_(scrh);
  const R cf2 = R(0.5) * (a2 + b2 + scrh);
  const R cs2 = div_(a2 * ca2, cf2);
  const R cf = sqrt_(cf2), cs = sqrt_(cs2), ca = sqrt_(ca2), a = sqrt_(a2);
  R alpha_f, alpha_s;
  if (cf == cs) {
    alpha_f = R(1); alpha_s = R(0);
  } else if (a <= cs) {
    alpha_f = R(0); alpha_s = R(1);
  } else if (cf <= a) {
    alpha_f = R(1); alpha_s = R(0);
  } else {
    scrh = div_(R(1), cf2 - cs2);
    alpha_f = (a2 - cs2) * scrh;
    alpha_s = (cf2 - a2) * scrh;
    alpha_f = max_ge(R(0), alpha_f);
    alpha_s = max_ge(R(0), alpha_s);
    alpha_f = sqrt_(alpha_f);
    alpha_s = sqrt_(alpha_s);
  }
  R beta_y, beta_z;
  if (two_comp) {
    beta_y = Btmag > R(1.e-9) ? (By >= R(0) ? R(1) : R(-1)) : R(1);
    beta_z = R(0);
  } else if (Btmag > R(1.e-9)) {
    beta_y = div_(By, Btmag);
    beta_z = div_(Bz, Btmag);
  } else {
    beta_z = beta_y = sqrt_1_2;
  }

  // fast wave u - c_f
  lambda[KFM] = u - cf;
  scrh = alpha_s * cs * sBx;
  const R beta_dv = R(0) + beta_y * dV[VXt] + beta_z * dV[VXb];
  const R beta_dB = R(0) + beta_y * dV[BXt] + beta_z * dV[BXb];
  const R beta_v = R(0) + beta_y * v + beta_z * w;
  Rc[RHO][KFM] = alpha_f;
  Rc[MXn][KFM] = alpha_f * lambda[KFM];
  Rc[MXt][KFM] = alpha_f * v + scrh * beta_y;
  Rc[MXb][KFM] = alpha_f * w + scrh * beta_z;
  Rc[BXt][KFM] = div_(alpha_s * a * beta_y, sqrt_rho);
  Rc[BXb][KFM] = div_(alpha_s * a * beta_z, sqrt_rho);
  Rc[ENG][KFM] = alpha_f * (Hgas - u * cf) + scrh * beta_v + div_(alpha_s * a * Btmag, sqrt_rho);
  eta[KFM] = alpha_f * (X * dV[RHO] + dV[PRS]) + rho * scrh * beta_dv - rho * alpha_f * cf * dV[VXn] +
             sqrt_rho * alpha_s * a * beta_dB;
  eta[KFM] *= div_(R(0.5), a2);

  // fast wave u + c_f
  lambda[KFP] = u + cf;
  Rc[RHO][KFP] = alpha_f;
  Rc[MXn][KFP] = alpha_f * lambda[KFP];
  Rc[MXt][KFP] = alpha_f * v - scrh * beta_y;
  Rc[MXb][KFP] = alpha_f * w - scrh * beta_z;
  Rc[BXt][KFP] = Rc[BXt][KFM];
  Rc[BXb][KFP] = Rc[BXb][KFM];
  Rc[ENG][KFP] = alpha_f * (Hgas + u * cf) - scrh * beta_v + div_(alpha_s * a * Btmag, sqrt_rho);
  eta[KFP] = alpha_f * (X * dV[RHO] + dV[PRS]) - rho * scrh * beta_dv + rho * alpha_f * cf * dV[VXn] +
             sqrt_rho * alpha_s * a * beta_dB;
  eta[KFP] *= div_(R(0.5), a2);

  // entropy wave u
  lambda[KEN] = u;
  Rc[RHO][KEN] = R(1);
  Rc[MXn][KEN] = u;
  Rc[MXt][KEN] = v;
  Rc[MXb][KEN] = w;
  Rc[ENG][KEN] = R(0.5) * vel2 + div_(gamma - R(2.0), g1) * X;
  eta[KEN] = div_((a2 - X) * dV[RHO] - dV[PRS], a2);

  // div.B wave u: zero jump with constrained transport
  lambda[KDB] = u;
  eta[KDB] = R(0);

  // slow wave u - c_s
  scrh = alpha_f * cf * sBx;
  lambda[KSM] = u - cs;
  Rc[RHO][KSM] = alpha_s;
  Rc[MXn][KSM] = alpha_s * lambda[KSM];
  Rc[MXt][KSM] = alpha_s * v - scrh * beta_y;
  Rc[MXb][KSM] = alpha_s * w - scrh * beta_z;
  Rc[BXt][KSM] = div_(-alpha_f * a * beta_y, sqrt_rho);
  Rc[BXb][KSM] = div_(-alpha_f * a * beta_z, sqrt_rho);
  Rc[ENG][KSM] = alpha_s * (Hgas - u * cs) - scrh * beta_v - div_(alpha_f * a * Btmag, sqrt_rho);
  eta[KSM] = alpha_s * (X * dV[RHO] + dV[PRS]) - rho * scrh * beta_dv - rho * alpha_s * cs * dV[VXn] -
             sqrt_rho * alpha_f * a * beta_dB;
  eta[KSM] *= div_(R(0.5), a2);

  // slow wave u + c_s
  lambda[KSP] = u + cs;
  Rc[RHO][KSP] = alpha_s;
  Rc[MXn][KSP] = alpha_s * lambda[KSP];
  Rc[MXt][KSP] = alpha_s * v + scrh * beta_y;
  Rc[MXb][KSP] = alpha_s * w + scrh * beta_z;
  Rc[BXt][KSP] = Rc[BXt][KSM];
  Rc[BXb][KSP] = Rc[BXb][KSM];
  Rc[ENG][KSP] = alpha_s * (Hgas + u * cs) + scrh * beta_v - div_(alpha_f * a * Btmag, sqrt_rho);
  eta[KSP] = alpha_s * (X * dV[RHO] + dV[PRS]) + rho * scrh * beta_dv + rho * alpha_s * cs * dV[VXn] -
             sqrt_rho * alpha_f * a * beta_dB;
  eta[KSP] *= div_(R(0.5), a2);

  // Alfven waves u -+ c_a
  lambda[KAM] = u - ca;
  Rc[MXt][KAM] = -rho * beta_z;
  Rc[MXb][KAM] = +rho * beta_y;
  Rc[BXt][KAM] = -sBx * sqrt_rho * beta_z;
  Rc[BXb][KAM] = sBx * sqrt_rho * beta_y;
  Rc[ENG][KAM] = -rho * (v * beta_z - w * beta_y);
  eta[KAM] = +beta_y * dV[VXb] - beta_z * dV[VXt] +
             div_(sBx, sqrt_rho) * (beta_y * dV[BXb] - beta_z * dV[BXt]);
  eta[KAM] *= R(0.5);
  lambda[KAP] = u + ca;
  Rc[MXt][KAP] = -Rc[MXt][KAM];
  Rc[MXb][KAP] = -Rc[MXb][KAM];
  Rc[BXt][KAP] = Rc[BXt][KAM];
  Rc[BXb][KAP] = Rc[BXb][KAM];
  Rc[ENG][KAP] = -Rc[ENG][KAM];
  eta[KAP] = -beta_y * dV[VXb] + beta_z * dV[VXt] +
             div_(sBx, sqrt_rho) * (beta_y * dV[BXb] - beta_z * dV[BXt]);
  eta[KAP] *= R(0.5);

  o.cmax = fabs(u) + cf;
  o.mach = fabs(div_(u, a));
#pragma unroll
  for (int k = 0; k < NV; k++) alambda[k] = fabs(lambda[k]);
  if (alambda[KFM] < R(0.5) * delta) alambda[KFM] = div_(lambda[KFM] * lambda[KFM], delta) + R(0.25) * delta;
  if (alambda[KFP] < R(0.5) * delta) alambda[KFP] = div_(lambda[KFP] * lambda[KFP], delta) + R(0.25) * delta;
  if (alambda[KSM] < R(0.5) * delta) alambda[KSM] = div_(lambda[KSM] * lambda[KSM], delta) + R(0.25) * delta;
  if (alambda[KSP] < R(0.5) * delta) alambda[KSP] = div_(lambda[KSP] * lambda[KSP], delta) + R(0.25) * delta;
  const int nwave = two_comp ? 6 : 8;
#pragma unroll
  for (int nv = 0; nv < NV; nv++) {
    if (!FULL && nv >= BX1 && nv <= BX3) continue;
    R s = R(0);
#pragma unroll
    for (int k = 0; k < NV; k++)
      if (k < nwave) s += alambda[k] * eta[k] * Rc[nv][k];
    o.f[nv] = R(0.5) * (fL[nv] + fR[nv] - s);
  }
  o.press = R(0.5) * (pL + pR);
  o.SL = lambda[KFM];
  o.SR = lambda[KFP];
}

constexpr int SOLVER_HLLD = 0;
constexpr int SOLVER_HLL = 1;
constexpr int SOLVER_HLLC = 2;
constexpr int SOLVER_TVDLF = 3;
constexpr int SOLVER_ROE = 4;

// two_comp: a 2D box, for the solvers whose COMPONENTS 2 build differs from
// the three-component one with zero z components (Roe)
template <int S, int D, bool FULL, typename R>
__device__ __forceinline__ void riemann(const R* vL, const R* vR, R gamma, R gmm1, RiemannOut<R>& o,
                                        bool two_comp = false, BranchRec* rec = nullptr)
{
  if constexpr (S == SOLVER_ROE) roe<D, FULL>(vL, vR, gamma, gmm1, o, two_comp);
  else if constexpr (S == SOLVER_HLL) hll<D, FULL>(vL, vR, gamma, gmm1, o);
  else if constexpr (S == SOLVER_HLLC) hllc<D, FULL>(vL, vR, gamma, gmm1, o);
  else if constexpr (S == SOLVER_TVDLF) tvdlf<D, FULL>(vL, vR, gamma, gmm1, o);
  else hlld<D, FULL>(vL, vR, gamma, gmm1, o, rec);   // branches recorded for HLLD only
}

// MC_LIM2 of the staggered-field slopes (MHD/CT/ct_stag_slopes.c:92-106).
__device__ __forceinline__ double mc_lim2(double dp, double dm, BranchRec* rec = nullptr)
{
  if (rb(rec, dp * dm < 0.0)) return 0.0;
  double dc = 0.5 * (dp + dm);
  double sc = 2.0 * (rb(rec, fabs(dp) < fabs(dm)) ? dp : dm);
  return rb(rec, fabs(dc) < fabs(sc)) ? dc : sc;
}

// Four-state HLL EMF at an edge (Londrillo & Del Zanna 2004, eq. 56; as in
// MHD/CT/ct_emf_average.c:285-289).
__device__ __forceinline__ double uct_hll(double a_xp, double a_xm, double a_yp, double a_ym,
                                          double eSW, double eSE, double eNE, double eNW,
                                          double bS, double bN, double bW, double bE)
{
  double e = a_xp * a_yp * eSW + a_xm * a_yp * eSE + a_xm * a_ym * eNE + a_xp * a_ym * eNW;
  e /= (a_xp + a_xm) * (a_yp + a_ym);
  e -= a_yp * a_ym * (bN - bS) / (a_yp + a_ym);
  e += a_xp * a_xm * (bE - bW) / (a_xp + a_xm);
  return e;
}

}  // namespace pgpu
