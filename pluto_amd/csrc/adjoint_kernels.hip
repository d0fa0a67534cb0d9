// Discrete adjoint of one RK2 step of the unsplit MHD stage (FP64, 3D,
// fixed dt), the reverse of the forward step in stage_kernels.hip:
//
//   stage 1   U*   = U0 + L(V, b),     b*  = b  - dt curl E(V, b),
//             V1   = ConsToPrim(U*, avg b*)
//   stage 2   U**  = U* + L(V1, b*),   b** = b* - dt curl E(V1, b*)
//   blend     U'   = (U0 + U**)/2,     b'  = (b + b**)/2,
//             V'   = ConsToPrim(U', avg b')          (rk_step.c:140-176)
//
// with U0 = PrimToCons(V), L the flux differences of update_stage.c /
// rhs.c (x, y, z; PLM-MC + HLLD), E the UCT_HLL edge EMF
// (ct_emf_average.c:179-352) and avg the face average of B
// (ct_field_average.c:64-66).  The adjoint step recomputes the forward
// step from the domain's state (the fused kernels), then runs the stages
// backwards.  Stencil adjoints that land on neighbouring cells are not
// scattered with atomics: the EMF's are written as per-edge partials and
// gathered per cell, the interfaces' are summed per cell in LDS (x rows) or
// in a per-thread ring while marching along y and z, so the adjoint is
// deterministic.  Floors are assumed inactive (smooth
// flows).

#include <hip/hip_runtime.h>

#include "domain.hpp"
#include "fused_kernels.hpp"
#include "mhd_adjoint.hpp"
#include "mhd_device.hpp"
#include "stage_kernels.hpp"

namespace pgpu {

namespace {

struct AdjCell {
  int i, j, k;
  long c;
};

__device__ __forceinline__ bool adj_cell(const Geo& g, AdjCell& p)
{
  p.i = blockIdx.x * blockDim.x + threadIdx.x;
  p.j = blockIdx.y * blockDim.y + threadIdx.y;
  p.k = blockIdx.z;
  if (p.i >= g.nx || p.j >= g.ny) return false;
  p.c = ((long)p.k * g.ny + p.j) * g.nx + p.i;
  return true;
}

// periodic neighbour at offset o (|o| <= 2) along axis A
template <int A>
__device__ __forceinline__ long nbr(const Geo& g, const AdjCell& p, int o)
{
  const int q = (A == 0 ? p.i : A == 1 ? p.j : p.k);
  const int nq = (A == 0 ? g.nx : A == 1 ? g.ny : g.nz);
  const long s = (A == 0 ? 1L : A == 1 ? (long)g.nx : (long)g.nx * g.ny);
  int qq = q + o;
  if (qq < 0) qq += nq;
  else if (qq >= nq) qq -= nq;
  return p.c + (long)(qq - q) * s;
}

__device__ __forceinline__ long nbr(const Geo& g, const AdjCell& p, int a, int o)
{
  return a == 0 ? nbr<0>(g, p, o) : a == 1 ? nbr<1>(g, p, o) : nbr<2>(g, p, o);
}

const dim3 kAdjBlock(64, 4);

dim3 adj_grid(const Geo& g)
{
  return dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
}

}  // namespace

// Adjoint work arrays (device), allocated on the first adjoint call.
struct AdjArrays {
  double* lamV[8];      // adjoint of the stage input primitives (cell B included)
  double* lamU[5];      // adjoint of the conservative accumulator
  double* lamU0[5];     // adjoint of U0 = PrimToCons(V^n) from the blend
  double* lamB[3];      // adjoint of the faces of the current stage
  double* lamB0[3];     // adjoint of b^n from the blend
  double* lamE[3];      // ex, ey, ez
  double* lamS[3][2];   // [direction][0: max(0,-SL), 1: max(0,SR)]
  double* lamDV[3][3];  // [direction][velocity component]
  double* lamBc[3];     // adjoint of the cell-centred field (temporary)
  double* VN[8];        // forward V^{n+1}
  double* bN[3];        // forward b^{n+1}
  double* ep[3][28];    // per-edge partials of the EMF adjoint, per edge family
  double* sw1[3][5];    // stage-1 sweep outputs kept for its reverse: SR, SL, dv[3]
};

// ---------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------

// Output side: V' = ConsToPrim(U', avg b'); lam_prim -> lamU (= lamU0 =
// half of the U' adjoint) and lamBc.
__global__ void __launch_bounds__(256)
k_adj_out(Geo g, AdjArrays A, const double* lam_prim)
{
  AdjCell p;
  if (!adj_cell(g, p)) return;
  double v[NV], lv[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    v[q] = A.VN[q][p.c];
    lv[q] = lam_prim[(long)q * g.n + p.c];
  }
  double au[5] = {0, 0, 0, 0, 0}, ab[3] = {lv[BX1], lv[BX2], lv[BX3]};
  cons2prim_vjp(v, lv, au, ab, g.gmm1);
#pragma unroll
  for (int q = 0; q < 5; q++) {
    A.lamU[q][p.c] = 0.5 * au[q];
    A.lamU0[q][p.c] = 0.5 * au[q];
  }
#pragma unroll
  for (int a = 0; a < 3; a++) A.lamBc[a][p.c] = ab[a];
#pragma unroll
  for (int q = 0; q < NV; q++) A.lamV[q][p.c] = 0.0;   // scatter target of stage 2
}

// lam_face + avg^T lamBc, halved into the stage-2 faces and the b^n part.
__global__ void __launch_bounds__(256)
k_adj_out_faces(Geo g, AdjArrays A, const double* lam_face)
{
  AdjCell p;
  if (!adj_cell(g, p)) return;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const long cp = nbr(g, p, a, 1);
    const double lb = lam_face[(long)a * g.n + p.c] + 0.5 * (A.lamBc[a][p.c] + A.lamBc[a][cp]);
    A.lamB[a][p.c] = 0.5 * lb;
    A.lamB0[a][p.c] = 0.5 * lb;
  }
}

// CT transpose: lamE from the face adjoint (ct_update.c:87-196 reversed).
__global__ void __launch_bounds__(256)
k_adj_ct(Geo g, AdjArrays A, const DevScalars* sc)
{
  AdjCell p;
  if (!adj_cell(g, p)) return;
  const double dt = sc->dt;
  const double dtx = dt / g.dx[0], dty = dt / g.dx[1], dtz = dt / g.dx[2];
  const long c = p.c;
  const long cip = nbr<0>(g, p, 1), cjp = nbr<1>(g, p, 1), ckp = nbr<2>(g, p, 1);
  const double *lbx = A.lamB[0], *lby = A.lamB[1], *lbz = A.lamB[2];
  A.lamE[2][c] = -dty * (lbx[c] - lbx[cjp]) + dtx * (lby[c] - lby[cip]);
  A.lamE[1][c] = dtz * (lbx[c] - lbx[ckp]) - dtx * (lbz[c] - lbz[cip]);
  A.lamE[0][c] = -dtz * (lby[c] - lby[ckp]) + dty * (lbz[c] - lbz[cjp]);
}

// ---------------------------------------------------------------------
// EMF adjoint in two passes without atomics, for the UCT_HLL edge families
// with axes (P, Q): Ez = <0,1>, Ex = <1,2>, Ey = <2,0> (the stencils of
// k_emf).  Pass 1 (per edge, for each family F) stores 28 partials:
//   0-7   speed adjoints: 0 SP.SR, 2 SP.SL, 4 SQ.SR, 6 SQ.SL, each going to
//         the cell the edge's max() selected (c or c+Q for SP, c or c+P
//         for SQ); the gather redoes the comparison, so slots 1, 3, 5, 7
//         are not stored
//   8-15  corner products (aq, ap) of SW (c), SE (c+P), NE (c+P+Q), NW (c+Q)
//   16-27 limited-face contributions: S to c, c+Q, c-Q; N to c+Q, c+2Q, c;
//         W to c, c+P, c-P; E to c+P, c+2P, c
// Pass 2 (per cell) gathers what the neighbouring edges sent to the cell.
// ---------------------------------------------------------------------

template <int P, int Q>
__device__ __forceinline__ void emf_edge_partials(const Geo& g, const AdjCell& p,
                                                  const StageIn& in, const Sweeps& sw,
                                                  const AdjArrays& A)
{
  constexpr int E = 3 - P - Q;
  constexpr int F = (P == 0 ? 0 : P == 1 ? 1 : 2);
  const long c = p.c;
  double o[28];
#pragma unroll
  for (int q = 0; q < 28; q++) o[q] = 0.0;
  const double le = A.lamE[E][c];
  if (le != 0.0) {
    const long cP = nbr<P>(g, p, 1), cQ = nbr<Q>(g, p, 1);
    const long cPQ = cP + (cQ - c);
    const SweepOut &SP = sw.d[P], &SQ = sw.d[Q];
    const double* vP = in.V[VX1 + P];
    const double* vQ = in.V[VX1 + Q];
    const double* bP = in.b[P];
    const double* bQ = in.b[Q];
    const bool sxp = SP.SR[c] >= SP.SR[cQ], sxm = SP.SL[c] >= SP.SL[cQ];
    const bool syp = SQ.SR[c] >= SQ.SR[cP], sym = SQ.SL[c] >= SQ.SL[cP];
    double x[12];
    x[0] = sxp ? SP.SR[c] : SP.SR[cQ];
    x[1] = sxm ? SP.SL[c] : SP.SL[cQ];
    x[2] = syp ? SQ.SR[c] : SQ.SR[cP];
    x[3] = sym ? SQ.SL[c] : SQ.SL[cP];
    const long cQm = nbr<Q>(g, p, -1), cPm = nbr<P>(g, p, -1);
    const long cQQ = nbr<Q>(g, p, 2), cPP = nbr<P>(g, p, 2);
    const double dpS = bP[cQ] - bP[c], dmS = bP[c] - bP[cQm];
    const double dpN = bP[cQQ] - bP[cQ], dmN = bP[cQ] - bP[c];
    const double dpW = bQ[cP] - bQ[c], dmW = bQ[c] - bQ[cPm];
    const double dpE = bQ[cPP] - bQ[cP], dmE = bQ[cP] - bQ[c];
    const double bS = bP[c] + 0.5 * mc_lim2(dpS, dmS);
    const double bW = bQ[c] + 0.5 * mc_lim2(dpW, dmW);
    const double bN = bP[cQ] - 0.5 * mc_lim2(dpN, dmN);
    const double bE = bQ[cP] - 0.5 * mc_lim2(dpE, dmE);
    const double qSW = vQ[c] + 0.5 * (SP.dv[Q][c] + SQ.dv[Q][c]);
    const double pSW = vP[c] + 0.5 * (SP.dv[P][c] + SQ.dv[P][c]);
    const double qSE = vQ[cP] - 0.5 * (SP.dv[Q][cP] - SQ.dv[Q][cP]);
    const double pSE = vP[cP] - 0.5 * (SP.dv[P][cP] - SQ.dv[P][cP]);
    const double qNE = vQ[cPQ] - 0.5 * (SP.dv[Q][cPQ] + SQ.dv[Q][cPQ]);
    const double pNE = vP[cPQ] - 0.5 * (SP.dv[P][cPQ] + SQ.dv[P][cPQ]);
    const double qNW = vQ[cQ] + 0.5 * (SP.dv[Q][cQ] - SQ.dv[Q][cQ]);
    const double pNW = vP[cQ] + 0.5 * (SP.dv[P][cQ] - SQ.dv[P][cQ]);
    x[4] = qSW * bS - pSW * bW;
    x[5] = qSE * bS - pSE * bE;
    x[6] = qNE * bN - pNE * bE;
    x[7] = qNW * bN - pNW * bW;
    x[8] = bS;
    x[9] = bN;
    x[10] = bW;
    x[11] = bE;
    double ax[12] = {0};
    uct_hll_vjp(x, le, ax);
    o[0] = ax[0];
    o[2] = ax[1];
    o[4] = ax[2];
    o[6] = ax[3];
    o[8] = ax[4] * bS;
    o[9] = -ax[4] * bW;
    o[10] = ax[5] * bS;
    o[11] = -ax[5] * bE;
    o[12] = ax[6] * bN;
    o[13] = -ax[6] * bE;
    o[14] = ax[7] * bN;
    o[15] = -ax[7] * bW;
    const double abS = ax[8] + ax[4] * qSW + ax[5] * qSE;
    const double abN = ax[9] + ax[6] * qNE + ax[7] * qNW;
    const double abW = ax[10] - ax[4] * pSW - ax[7] * pNW;
    const double abE = ax[11] - ax[5] * pSE - ax[6] * pNE;
    double adp, adm;
    adp = adm = 0.0;
    mc_lim2_vjp(dpS, dmS, 0.5 * abS, adp, adm);
    o[16] = abS - adp + adm;
    o[17] = adp;
    o[18] = -adm;
    adp = adm = 0.0;
    mc_lim2_vjp(dpN, dmN, -0.5 * abN, adp, adm);
    o[19] = abN - adp + adm;
    o[20] = adp;
    o[21] = -adm;
    adp = adm = 0.0;
    mc_lim2_vjp(dpW, dmW, 0.5 * abW, adp, adm);
    o[22] = abW - adp + adm;
    o[23] = adp;
    o[24] = -adm;
    adp = adm = 0.0;
    mc_lim2_vjp(dpE, dmE, -0.5 * abE, adp, adm);
    o[25] = abE - adp + adm;
    o[26] = adp;
    o[27] = -adm;
  }
#pragma unroll
  for (int q = 0; q < 28; q++)
    if (q >= 8 || !(q & 1)) A.ep[F][q][c] = o[q];
}

__global__ void __launch_bounds__(256)
k_adj_emf_partials(Geo g, StageIn in, Sweeps sw, AdjArrays A)
{
  AdjCell p;
  if (!adj_cell(g, p)) return;
  emf_edge_partials<0, 1>(g, p, in, sw, A);   // Ez
  emf_edge_partials<1, 2>(g, p, in, sw, A);   // Ex
  emf_edge_partials<2, 0>(g, p, in, sw, A);   // Ey
}

// periodic neighbour of p at offsets (di, dj, dk), each in [-2, 2]
__device__ __forceinline__ long nbr3(const Geo& g, const AdjCell& p, int di, int dj, int dk)
{
  int i = p.i + di, j = p.j + dj, k = p.k + dk;
  i = i < 0 ? i + g.nx : (i >= g.nx ? i - g.nx : i);
  j = j < 0 ? j + g.ny : (j >= g.ny ? j - g.ny : j);
  k = k < 0 ? k + g.nz : (k >= g.nz ? k - g.nz : k);
  return ((long)k * g.ny + j) * g.nx + i;
}

template <int P, int Q>
__device__ __forceinline__ void emf_gather_family(const Geo& g, const AdjCell& p,
                                                  const AdjArrays& A, const Sweeps& sw,
                                                  double (&sS)[3][2],
                                                  double (&sdv)[3][3], double (&sv)[3],
                                                  double (&sb)[3])
{
  constexpr int F = (P == 0 ? 0 : P == 1 ? 1 : 2);
  auto at = [&](int slot, int dp, int dq) {
    int d[3] = {0, 0, 0};
    d[P] += dp;
    d[Q] += dq;
    return A.ep[F][slot][nbr3(g, p, d[0], d[1], d[2])];
  };
  const long c = p.c;
  const double* const* e = A.ep[F];
  // speeds: the max() of edge c picked cell c, that of edge c-Q (c-P)
  // picked its upper cell c (the comparisons of emf_edge_partials)
  {
    const SweepOut &SP = sw.d[P], &SQ = sw.d[Q];
    int dq[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
    dq[Q] = 1;
    dp[P] = 1;
    const long cQ = nbr3(g, p, dq[0], dq[1], dq[2]), cQm = nbr3(g, p, -dq[0], -dq[1], -dq[2]);
    const long cP = nbr3(g, p, dp[0], dp[1], dp[2]), cPm = nbr3(g, p, -dp[0], -dp[1], -dp[2]);
    sS[P][1] += (SP.SR[c] >= SP.SR[cQ] ? e[0][c] : 0.0) + (SP.SR[cQm] >= SP.SR[c] ? 0.0 : e[0][cQm]);
    sS[P][0] += (SP.SL[c] >= SP.SL[cQ] ? e[2][c] : 0.0) + (SP.SL[cQm] >= SP.SL[c] ? 0.0 : e[2][cQm]);
    sS[Q][1] += (SQ.SR[c] >= SQ.SR[cP] ? e[4][c] : 0.0) + (SQ.SR[cPm] >= SQ.SR[c] ? 0.0 : e[4][cPm]);
    sS[Q][0] += (SQ.SL[c] >= SQ.SL[cP] ? e[6][c] : 0.0) + (SQ.SL[cPm] >= SQ.SL[c] ? 0.0 : e[6][cPm]);
  }
  // corners: c is SW of edge c, SE of c-P, NE of c-P-Q, NW of c-Q
  const double aSW = e[8][c], pSWv = e[9][c];
  const double aSE = at(10, -1, 0), pSEv = at(11, -1, 0);
  const double aNE = at(12, -1, -1), pNEv = at(13, -1, -1);
  const double aNW = at(14, 0, -1), pNWv = at(15, 0, -1);
  sv[Q] += aSW + aSE + aNE + aNW;
  sv[P] += pSWv + pSEv + pNEv + pNWv;
  sdv[P][Q] += 0.5 * aSW - 0.5 * aSE - 0.5 * aNE + 0.5 * aNW;
  sdv[Q][Q] += 0.5 * aSW + 0.5 * aSE - 0.5 * aNE - 0.5 * aNW;
  sdv[P][P] += 0.5 * pSWv - 0.5 * pSEv - 0.5 * pNEv + 0.5 * pNWv;
  sdv[Q][P] += 0.5 * pSWv + 0.5 * pSEv - 0.5 * pNEv - 0.5 * pNWv;
  // faces
  sb[P] += e[16][c] + at(17, 0, -1) + at(18, 0, 1) + at(19, 0, -1) + at(20, 0, -2) + e[21][c];
  sb[Q] += e[22][c] + at(23, -1, 0) + at(24, 1, 0) + at(25, -1, 0) + at(26, -2, 0) + e[27][c];
}

__global__ void __launch_bounds__(256) k_adj_emf_gather(Geo g, Sweeps sw, AdjArrays A)
{
  AdjCell p;
  if (!adj_cell(g, p)) return;
  double sS[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  double sdv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double sv[3] = {0, 0, 0}, sb[3] = {0, 0, 0};
  emf_gather_family<0, 1>(g, p, A, sw, sS, sdv, sv, sb);
  emf_gather_family<1, 2>(g, p, A, sw, sS, sdv, sv, sb);
  emf_gather_family<2, 0>(g, p, A, sw, sS, sdv, sv, sb);
  const long c = p.c;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    A.lamS[d][0][c] = sS[d][0];
    A.lamS[d][1][c] = sS[d][1];
#pragma unroll
    for (int k = 0; k < 3; k++) A.lamDV[d][k][c] = sdv[d][k];
    A.lamV[VX1 + d][c] += sv[d];
    A.lamB[d][c] += sb[d];
  }
}

// Adjoint of the interface q+1/2 along D (q = cell p): U += -dt/dx
// (F_{q+1/2} - F_{q-1/2}) and the pressure term, the UCT_HLL face speeds
// and the velocity slopes of cell q, through HLLD and the MC-limited PLM.
// part[4 s + k] receives what variable slot s (BN skipped) of the cell at
// offset k - 1 from q along D gets; the face adjoint of the interface
// itself is added to lamB when own is set.
template <int D>
__device__ __forceinline__ void adj_interface(const Geo& g, const StageIn& in, const AdjArrays& A,
                                              double dtdx, const AdjCell& p, bool own,
                                              double part[28])
{
  constexpr int BN = BX1 + Dir<D>::n;
  const long c = p.c;
  const long cm1 = nbr<D>(g, p, -1), cp1 = nbr<D>(g, p, 1), cp2 = nbr<D>(g, p, 2);
  double vm[NV], v0[NV], vp[NV], vq[NV], vL[NV], vR[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    const double* a = in.V[q];
    vm[q] = a[cm1];
    v0[q] = a[c];
    vp[q] = a[cp1];
    vq[q] = a[cp2];
    vL[q] = v0[q] + mc_slope(vm[q], v0[q], vp[q]) * 0.5;
    vR[q] = vp[q] - mc_slope(v0[q], vp[q], vq[q]) * 0.5;
  }
  const double bface = in.b[D][c];
  vL[BN] = bface;
  vR[BN] = bface;

  HlldSeed s;
#pragma unroll
  for (int q = 0; q < NV; q++) s.f[q] = 0.0;
  s.f[RHO] = dtdx * (A.lamU[0][cp1] - A.lamU[0][c]);
  s.f[VX1] = dtdx * (A.lamU[1][cp1] - A.lamU[1][c]);
  s.f[VX2] = dtdx * (A.lamU[2][cp1] - A.lamU[2][c]);
  s.f[VX3] = dtdx * (A.lamU[3][cp1] - A.lamU[3][c]);
  s.f[ENG] = dtdx * (A.lamU[4][cp1] - A.lamU[4][c]);
  s.press = dtdx * (A.lamU[1 + D][cp1] - A.lamU[1 + D][c]);
  // stored speeds max(0, -SL), max(0, SR): derivative on the active side
  double slmin, slmax, srmin, srmax;
  fast_speeds<D>(vL, g.gamma, slmin, slmax);
  fast_speeds<D>(vR, g.gamma, srmin, srmax);
  const double SL = min_le(slmin, srmin), SR = max_ge(slmax, srmax);
  s.SL = (0.0 >= -SL) ? 0.0 : -A.lamS[D][0][c];
  s.SR = (0.0 >= SR) ? 0.0 : A.lamS[D][1][c];

  double aL[NV] = {0}, aR[NV] = {0};
  hlld_vjp<D>(vL, vR, g.gamma, g.gmm1, s, aL, aR);
  if (own) A.lamB[D][c] += aL[BN] + aR[BN];

#pragma unroll
  for (int q = 0; q < NV; q++) {
    if (q == BN) continue;
    double a_s0 = 0.5 * aL[q];
    if (q >= VX1 && q <= VX3) a_s0 += A.lamDV[D][q - VX1][c];
    const double a_s1 = -0.5 * aR[q];
    double am = 0.0, a0 = aL[q], ap = aR[q], aq = 0.0;
    mc_slope_vjp(vm[q], v0[q], vp[q], a_s0, am, a0, ap);
    mc_slope_vjp(v0[q], vp[q], vq[q], a_s1, a0, ap, aq);
    const int sl = 4 * (q < BN ? q : q - 1);
    part[sl + 0] = am;
    part[sl + 1] = a0;
    part[sl + 2] = ap;
    part[sl + 3] = aq;
  }
}

// The x interfaces of one box row (one workgroup, thread i = cell i, nx <=
// 256): each thread solves the interface i+1/2, the 28 partials go through
// LDS, and cell i sums what the interfaces i+3/2, i+1/2, i-1/2, i-3/2 sent
// it into lamV (periodic along the row).
__global__ void __launch_bounds__(256) k_adj_flux_x_rows(Geo g, StageIn in, AdjArrays A,
                                                         const DevScalars* sc)
{
  extern __shared__ double sm[];   // [28][blockDim.x]
  const int nt = blockDim.x, i = threadIdx.x;
  AdjCell p;
  p.i = i;
  p.j = blockIdx.x;
  p.k = blockIdx.y;
  p.c = ((long)p.k * g.ny + p.j) * g.nx + i;
  if (i < g.nx) {
    double part[28];
    adj_interface<0>(g, in, A, sc->dt / g.dx[0], p, true, part);
#pragma unroll
    for (int m = 0; m < 28; m++) sm[m * nt + i] = part[m];
  }
  __syncthreads();
  if (i >= g.nx) return;
  const int n1 = i + 1 < g.nx ? i + 1 : i + 1 - g.nx;
  const int m1 = i >= 1 ? i - 1 : i - 1 + g.nx;
  const int m2 = i >= 2 ? i - 2 : i - 2 + g.nx;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    if (q == BX1) continue;
    const int sl = 4 * (q < BX1 ? q : q - 1);
    A.lamV[q][p.c] += sm[(sl + 0) * nt + n1] + sm[(sl + 1) * nt + i] +
                      sm[(sl + 2) * nt + m1] + sm[(sl + 3) * nt + m2];
  }
}

// The interfaces along D of one line segment per thread: lanes run along
// the first transverse axis (x for D = 1, 2; y for D = 0), blockIdx.y picks
// the line in the other one, blockIdx.z the segment [q0, q0 + len) of
// cells along D.  The thread solves the interfaces q+1/2 for q = q0-2 ..
// q0+len in order (the three beyond the segment, q0-2, q0-1 and q0+len, are
// solved again by the neighbouring segments), adds each one's partials into a ring of four
// cells kept in LDS, and after interface q adds the now complete cell q-1
// to lamV.  Only the interfaces q0..q0+len-1 add their face adjoint.
template <int D>
__global__ void __launch_bounds__(256) k_adj_flux_march(Geo g, StageIn in, AdjArrays A,
                                                        const DevScalars* sc, int len)
{
  constexpr int BN = BX1 + Dir<D>::n;
  __shared__ double ring[28][256];   // [4 * slot + cell (mod 4)][thread]
  const int t = threadIdx.y * blockDim.x + threadIdx.x;
  const int nD = D == 0 ? g.nx : D == 1 ? g.ny : g.nz;
  const int l0 = blockIdx.x * blockDim.x + threadIdx.x;           // first transverse axis
  const int l1 = blockIdx.y * blockDim.y + threadIdx.y;           // second one
  const int n0 = D == 0 ? g.ny : g.nx, n1 = D == 2 ? g.ny : g.nz;
  if (l0 >= n0 || l1 >= n1) return;
  const int q0 = blockIdx.z * len;
  const int q1 = min(q0 + len, nD);
#pragma unroll
  for (int m = 0; m < 28; m++) ring[m][t] = 0.0;
  const double dtdx = sc->dt / g.dx[D];
  for (int q = q0 - 2; q <= q1; q++) {
    const int qw = q < 0 ? q + nD : q >= nD ? q - nD : q;
    AdjCell p;
    p.i = D == 0 ? qw : l0;
    p.j = D == 1 ? qw : D == 0 ? l0 : l1;
    p.k = D == 2 ? qw : l1;
    p.c = ((long)p.k * g.ny + p.j) * g.nx + p.i;
    double part[28];
    adj_interface<D>(g, in, A, dtdx, p, q >= q0 && q < q1, part);
    // targets q-1 .. q+2 (k = 0..3) sit in ring slots (q - 1 + k) mod 4
#pragma unroll
    for (int s = 0; s < 7; s++)
#pragma unroll
      for (int k = 0; k < 4; k++) ring[4 * s + ((q - 1 + k) & 3)][t] += part[4 * s + k];
    // cell q-1 is complete: its last interface was q+1/2
    const int done = (q - 1) & 3;
    if (q - 1 >= q0) {
      const int cw = q - 1 < 0 ? q - 1 + nD : q - 1;
      const long cc = p.c + (long)(cw - qw) * (D == 0 ? 1L : D == 1 ? (long)g.nx : (long)g.nx * g.ny);
#pragma unroll
      for (int v = 0; v < NV; v++) {
        if (v == BN) continue;
        const int s = v < BN ? v : v - 1;
        A.lamV[v][cc] += ring[4 * s + done][t];
      }
    }
#pragma unroll
    for (int s = 0; s < 7; s++) ring[4 * s + done][t] = 0.0;
  }
}

// Between the stages: V1 = ConsToPrim(U*, avg b*).  lamU += (dV1/dU*)^T lamV,
// lamBc = field part; lamV is cleared for the stage-1 pass.
__global__ void __launch_bounds__(256)
k_adj_mid(Geo g, AdjArrays A, StageIn in)
{
  AdjCell p;
  if (!adj_cell(g, p)) return;
  double v[NV], lv[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    v[q] = in.V[q][p.c];
    lv[q] = A.lamV[q][p.c];
    A.lamV[q][p.c] = 0.0;
  }
  double au[5] = {0, 0, 0, 0, 0}, ab[3] = {lv[BX1], lv[BX2], lv[BX3]};
  cons2prim_vjp(v, lv, au, ab, g.gmm1);
#pragma unroll
  for (int q = 0; q < 5; q++) A.lamU[q][p.c] += au[q];
#pragma unroll
  for (int a = 0; a < 3; a++) A.lamBc[a][p.c] = ab[a];
}

// lamB += avg^T lamBc
__global__ void __launch_bounds__(256) k_adj_avgT(Geo g, AdjArrays A)
{
  AdjCell p;
  if (!adj_cell(g, p)) return;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const long cp = nbr(g, p, a, 1);
    A.lamB[a][p.c] += 0.5 * (A.lamBc[a][p.c] + A.lamBc[a][cp]);
  }
}

// Input side: U0 = PrimToCons(V^n) takes lamU (stage 1) + lamU0 (blend);
// writes the primitive adjoint (field components folded into lamBc).
__global__ void __launch_bounds__(256)
k_adj_in(Geo g, AdjArrays A, StageIn in, double* lam_prim)
{
  AdjCell p;
  if (!adj_cell(g, p)) return;
  double v[NV], lv[NV], au[NV] = {0};
#pragma unroll
  for (int q = 0; q < NV; q++) {
    v[q] = in.V[q][p.c];
    lv[q] = A.lamV[q][p.c];
  }
  au[RHO] = A.lamU[0][p.c] + A.lamU0[0][p.c];
  au[VX1] = A.lamU[1][p.c] + A.lamU0[1][p.c];
  au[VX2] = A.lamU[2][p.c] + A.lamU0[2][p.c];
  au[VX3] = A.lamU[3][p.c] + A.lamU0[3][p.c];
  au[ENG] = A.lamU[4][p.c] + A.lamU0[4][p.c];
  prim2cons_vjp(v, au, lv, g.gmm1);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    A.lamBc[a][p.c] = lv[BX1 + a];
    lv[BX1 + a] = 0.0;
  }
#pragma unroll
  for (int q = 0; q < NV; q++) lam_prim[(long)q * g.n + p.c] = lv[q];
}

__global__ void __launch_bounds__(256)
k_adj_faces_final(Geo g, AdjArrays A, double* lam_face)
{
  AdjCell p;
  if (!adj_cell(g, p)) return;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const long cp = nbr(g, p, a, 1);
    lam_face[(long)a * g.n + p.c] =
        A.lamB[a][p.c] + 0.5 * (A.lamBc[a][p.c] + A.lamBc[a][cp]) + A.lamB0[a][p.c];
  }
}

// Cell-centred field as the face average (after a device-side state copy).
__global__ void __launch_bounds__(256) k_cell_field(Geo g, Prims V, FaceIn b)
{
  AdjCell p;
  if (!adj_cell(g, p)) return;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (a == 2 && g.dims != 3) {
      V.v[BX3][p.c] = 0.0;
      continue;
    }
    const long cm = nbr(g, p, a, -1);
    V.v[BX1 + a][p.c] = 0.5 * (b.b[a][p.c] + b.b[a][cm]);
  }
}

// Segment length of the interface march along D: the longest of 64, 32, 16
// that still gives one workgroup per CU (the march runs one wave per SIMD,
// one workgroup per CU at a time); each segment solves the three interfaces
// beyond its cells again.
static int adj_march_len(const Geo& g, int D)
{
  const int nD = D == 0 ? g.nx : D == 1 ? g.ny : g.nz;
  const long lines = D == 0 ? (long)((g.ny + 63) / 64) * ((g.nz + 3) / 4)
                            : (long)((g.nx + 63) / 64) * (((D == 1 ? g.nz : g.ny) + 3) / 4);
  int len = 64;
  while (len > 16 && lines * ((nD + len - 1) / len) < 256) len /= 2;
  return len;
}

// One stage in reverse: CT transpose, EMF adjoint, flux adjoints.
static void adj_stage(const Geo& g, const StageIn& in, const Sweeps& sw, const AdjArrays& A,
                      const DevScalars* sc, hipStream_t s)
{
  const dim3 gr = adj_grid(g);
  hipLaunchKernelGGL(k_adj_ct, gr, kAdjBlock, 0, s, g, A, sc);
  hipLaunchKernelGGL(k_adj_emf_partials, gr, kAdjBlock, 0, s, g, in, sw, A);
  hipLaunchKernelGGL(k_adj_emf_gather, gr, kAdjBlock, 0, s, g, sw, A);
  // x: whole rows through LDS where they fit a workgroup, else marching
  if (g.nx <= 256) {
    const int nt = (g.nx + 63) / 64 * 64;
    hipLaunchKernelGGL(k_adj_flux_x_rows, dim3(g.ny, g.nz), dim3(nt), 28 * nt * sizeof(double), s,
                       g, in, A, sc);
  } else {
    const int len = adj_march_len(g, 0);
    hipLaunchKernelGGL(k_adj_flux_march<0>, dim3((g.ny + 63) / 64, (g.nz + 3) / 4, (g.nx + len - 1) / len),
                       kAdjBlock, 0, s, g, in, A, sc, len);
  }
  const int ly = adj_march_len(g, 1), lz = adj_march_len(g, 2);
  hipLaunchKernelGGL(k_adj_flux_march<1>, dim3((g.nx + 63) / 64, (g.nz + 3) / 4, (g.ny + ly - 1) / ly),
                     kAdjBlock, 0, s, g, in, A, sc, ly);
  hipLaunchKernelGGL(k_adj_flux_march<2>, dim3((g.nx + 63) / 64, (g.ny + 3) / 4, (g.nz + lz - 1) / lz),
                     kAdjBlock, 0, s, g, in, A, sc, lz);
}

}  // namespace pgpu

using namespace pgpu;

struct pgpu_adjoint_state {
  AdjArrays a;
  void* block = nullptr;
  size_t bytes = 0;
};

static int adj_alloc(pgpu_domain* d)
{
  if (d->adj) return PGPU_OK;
  const size_t per = (size_t)d->g.n * sizeof(double);
  const int narr = 8 + 5 + 5 + 3 + 3 + 3 + 6 + 9 + 3 + 8 + 3 + 84 + 15;
  pgpu_adjoint_state* st = new pgpu_adjoint_state;
  st->bytes = per * narr;
  if (hipMalloc(&st->block, st->bytes) != hipSuccess) {
    const size_t b = st->bytes;
    delete st;
    return fail(PGPU_ENOMEM, "hipMalloc of %zu bytes for the adjoint failed", b);
  }
  char* cur = (char*)st->block;
  auto take = [&]() {
    double* p = (double*)cur;
    cur += per;
    return p;
  };
  AdjArrays& A = st->a;
  for (int q = 0; q < 8; q++) A.lamV[q] = take();
  for (int q = 0; q < 5; q++) A.lamU[q] = take();
  for (int q = 0; q < 5; q++) A.lamU0[q] = take();
  for (int a = 0; a < 3; a++) A.lamB[a] = take();
  for (int a = 0; a < 3; a++) A.lamB0[a] = take();
  for (int a = 0; a < 3; a++) A.lamE[a] = take();
  for (int dd = 0; dd < 3; dd++)
    for (int k = 0; k < 2; k++) A.lamS[dd][k] = take();
  for (int dd = 0; dd < 3; dd++)
    for (int k = 0; k < 3; k++) A.lamDV[dd][k] = take();
  for (int a = 0; a < 3; a++) A.lamBc[a] = take();
  for (int q = 0; q < 8; q++) A.VN[q] = take();
  for (int a = 0; a < 3; a++) A.bN[a] = take();
  for (int f = 0; f < 3; f++)
    for (int q = 0; q < 28; q++) A.ep[f][q] = take();
  for (int dd = 0; dd < 3; dd++)
    for (int q = 0; q < 5; q++) A.sw1[dd][q] = take();
  d->adj = st;
  return PGPU_OK;
}

void pgpu_adjoint_free(pgpu_domain* d)
{
  if (!d || !d->adj) return;
  (void)hipFree(d->adj->block);
  delete d->adj;
  d->adj = nullptr;
}

extern "C" int pgpu_set_state_device(pgpu_domain* d, const double* prim8, const double* face3)
{
  if (!d || !prim8 || !face3) return fail(PGPU_EINVAL, "NULL argument");
  if (d->g.ghost) return fail(PGPU_EINVAL, "a bounded box takes its state through pgpu_set_state");
  const Geo& g = d->g;
  const size_t nb = (size_t)g.n * sizeof(double);
  for (int q = 0; q < 8; q++)
    if (hipMemcpyAsync(d->V[q], prim8 + (size_t)q * g.n, nb, hipMemcpyDeviceToDevice,
                       d->stream) != hipSuccess)
      return fail(PGPU_EDEVICE, "device state copy failed");
  for (int a = 0; a < 3; a++)
    if (hipMemcpyAsync(d->bA[a], face3 + (size_t)a * g.n, nb, hipMemcpyDeviceToDevice,
                       d->stream) != hipSuccess)
      return fail(PGPU_EDEVICE, "device state copy failed");
  Prims V;
  FaceIn b;
  for (int q = 0; q < 8; q++) V.v[q] = d->V[q];
  for (int a = 0; a < 3; a++) b.b[a] = d->bA[a];
  hipLaunchKernelGGL(k_cell_field, adj_grid(g), kAdjBlock, 0, d->stream, g, V, b);
  if (hipGetLastError() != hipSuccess) return fail(PGPU_EDEVICE, "cell field kernel failed");
  return PGPU_OK;
}

extern "C" int pgpu_get_state_device(pgpu_domain* d, double* prim8, double* face3)
{
  if (!d || !prim8 || !face3) return fail(PGPU_EINVAL, "NULL argument");
  if (d->g.ghost) return fail(PGPU_EINVAL, "a bounded box returns its state through pgpu_get_state");
  const Geo& g = d->g;
  const size_t nb = (size_t)g.n * sizeof(double);
  for (int q = 0; q < 8; q++)
    if (hipMemcpyAsync(prim8 + (size_t)q * g.n, d->V[q], nb, hipMemcpyDeviceToDevice,
                       d->stream) != hipSuccess)
      return fail(PGPU_EDEVICE, "device state copy failed");
  for (int a = 0; a < 3; a++)
    if (hipMemcpyAsync(face3 + (size_t)a * g.n, d->bA[a], nb, hipMemcpyDeviceToDevice,
                       d->stream) != hipSuccess)
      return fail(PGPU_EDEVICE, "device state copy failed");
  return PGPU_OK;
}

// What the reverse of one step reads from its forward: the step's input
// (V^n, b^n), the stage-1 output (V1, b*), the face speeds and velocity
// slopes of both stages, and the output V^{n+1}.  Held in the domain's and
// the adjoint's own arrays (pgpu_adjoint_step) or in a caller's tape of
// PGPU_ADJOINT_TAPE_ARRAYS arrays of n doubles, in this order.
struct StepTape {
  double* V[8];
  double* b[3];
  double* V1[8];
  double* b1[3];
  double* sw[2][3][5];   // [stage][direction][SR, SL, dv0, dv1, dv2]
  double* VN[8];
};

static StepTape tape_view(double* base, long n)
{
  StepTape T;
  double* cur = base;
  auto take = [&]() {
    double* p = cur;
    cur += n;
    return p;
  };
  for (int q = 0; q < 8; q++) T.V[q] = take();
  for (int a = 0; a < 3; a++) T.b[a] = take();
  for (int q = 0; q < 8; q++) T.V1[q] = take();
  for (int a = 0; a < 3; a++) T.b1[a] = take();
  for (int st = 0; st < 2; st++)
    for (int dd = 0; dd < 3; dd++)
      for (int q = 0; q < 5; q++) T.sw[st][dd][q] = take();
  for (int q = 0; q < 8; q++) T.VN[q] = take();
  return T;
}

static Sweeps tape_sweeps(const pgpu_domain* d, const StepTape& T, int st)
{
  Sweeps sw = d->sw;   // flux arrays of the modular sweeps: unused here
  for (int dd = 0; dd < 3; dd++) {
    sw.d[dd].SR = T.sw[st][dd][0];
    sw.d[dd].SL = T.sw[st][dd][1];
    for (int c = 0; c < 3; c++) sw.d[dd].dv[c] = T.sw[st][dd][2 + c];
  }
  return sw;
}

static int adj_checks(const pgpu_domain* d)
{
  const Geo& g = d->g;
  if (g.dims != 3) return fail(PGPU_EINVAL, "the adjoint step needs DIMENSIONS 3");
  if (g.solver != SOLVER_HLLD) return fail(PGPU_EINVAL, "the adjoint step is the HLLD solver's");
  if (g.emf_avg != EMF_UCT_HLL) return fail(PGPU_EINVAL, "the adjoint step uses the UCT_HLL EMF average");
  if (g.ghost) return fail(PGPU_EINVAL, "the adjoint step is for periodic boxes");
  if (!d->fixed_dt) return fail(PGPU_EINVAL, "the adjoint step needs a fixed dt (pgpu_set_dt)");
  if (g.nx < 3 || g.ny < 3 || g.nz < 3)
    return fail(PGPU_EINVAL, "the adjoint step needs at least 3 cells per direction");
  return PGPU_OK;
}

// The forward step from T.V, T.b (the fused FP64 kernels, bit-identical to
// the step pgpu_run takes): stage 1 into (T.V1, T.b1), stage 2 into
// (T.VN, outB), face data into T.sw.  U, the edges, C_dt and the flags are
// the domain's scratch.
static int adj_forward(pgpu_domain* d, const StepTape& T, double* const outB[3])
{
  const Geo& g = d->g;
  hipStream_t s = d->stream;
  StageIn in1, in2, vn;
  for (int q = 0; q < 8; q++) {
    in1.V[q] = T.V[q];
    in2.V[q] = T.V1[q];
    vn.V[q] = T.V[q];
  }
  for (int a = 0; a < 3; a++) {
    in1.b[a] = T.b[a];
    in2.b[a] = T.b1[a];
    vn.b[a] = T.b[a];
  }
  Cons U;
  for (int q = 0; q < 5; q++) U.u[q] = d->U[q];
  Prims V1, VN;
  for (int q = 0; q < 8; q++) {
    V1.v[q] = T.V1[q];
    VN.v[q] = T.VN[q];
  }
  FaceIn fA, fB;
  FaceOut oB, oN;
  for (int a = 0; a < 3; a++) {
    fA.b[a] = T.b[a];
    fB.b[a] = T.b1[a];
    oB.b[a] = T.b1[a];
    oN.b[a] = outB[a];
  }
  const Sweeps sw1 = tape_sweeps(d, T, 0), sw2 = tape_sweeps(d, T, 1);
  FaceData fd1, fd2;
  for (int dd = 0; dd < 3; dd++) {
    fd1.SR[dd] = sw1.d[dd].SR;
    fd1.SL[dd] = sw1.d[dd].SL;
    fd2.SR[dd] = sw2.d[dd].SR;
    fd2.SL[dd] = sw2.d[dd].SL;
    for (int c = 0; c < 3; c++) {
      fd1.dv[dd][c] = sw1.d[dd].dv[c];
      fd2.dv[dd][c] = sw2.d[dd].dv[c];
    }
  }
  int rc = launch_zero_flags(g, d->flag, s);
  if (!rc) rc = launch_flux_fused(g, 1, false, in1, U, fd1, d->cdt, d->sc, s, nullptr);
  if (!rc) rc = launch_emf(g, in1, sw1, d->e, s, nullptr);
  if (!rc) rc = launch_ct_c2p(g, 1, d->e, fA, fA, oB, U, vn, V1, d->flag, d->sc, s, nullptr);
  if (!rc) rc = launch_flux_fused(g, 2, false, in2, U, fd2, d->cdt, d->sc, s, nullptr);
  if (!rc) rc = launch_emf(g, in2, sw2, d->e, s, nullptr);
  if (!rc) rc = launch_ct_c2p(g, 2, d->e, fB, fA, oN, U, vn, VN, d->flag, d->sc, s, nullptr);
  if (rc) return fail(PGPU_EDEVICE, "forward recomputation failed: %s", hipGetErrorString((hipError_t)rc));
  return PGPU_OK;
}

// The reverse of the step T holds: output side, stage 2, stage 1, input side.
static int adj_reverse(pgpu_domain* d, const StepTape& T, double* lam_prim, double* lam_face)
{
  const Geo& g = d->g;
  hipStream_t s = d->stream;
  AdjArrays A = d->adj->a;
  for (int q = 0; q < 8; q++) A.VN[q] = T.VN[q];
  StageIn in1, in2;
  for (int q = 0; q < 8; q++) {
    in1.V[q] = T.V[q];
    in2.V[q] = T.V1[q];
  }
  for (int a = 0; a < 3; a++) {
    in1.b[a] = T.b[a];
    in2.b[a] = T.b1[a];
  }
  const Sweeps sw1 = tape_sweeps(d, T, 0), sw2 = tape_sweeps(d, T, 1);
  const dim3 gr = adj_grid(g);
  // output side and blend
  hipLaunchKernelGGL(k_adj_out, gr, kAdjBlock, 0, s, g, A, (const double*)lam_prim);
  hipLaunchKernelGGL(k_adj_out_faces, gr, kAdjBlock, 0, s, g, A, (const double*)lam_face);
  // stage 2
  adj_stage(g, in2, sw2, A, d->sc, s);
  hipLaunchKernelGGL(k_adj_mid, gr, kAdjBlock, 0, s, g, A, in2);
  hipLaunchKernelGGL(k_adj_avgT, gr, kAdjBlock, 0, s, g, A);
  // stage 1
  adj_stage(g, in1, sw1, A, d->sc, s);
  hipLaunchKernelGGL(k_adj_in, gr, kAdjBlock, 0, s, g, A, in1, lam_prim);
  hipLaunchKernelGGL(k_adj_faces_final, gr, kAdjBlock, 0, s, g, A, lam_face);
  if (hipGetLastError() != hipSuccess) return fail(PGPU_EDEVICE, "adjoint kernel launch failed");
  return PGPU_OK;
}

extern "C" int pgpu_adjoint_step(pgpu_domain* d, double* lam_prim, double* lam_face)
{
  if (!d || !lam_prim || !lam_face) return fail(PGPU_EINVAL, "NULL argument");
  int rc = adj_checks(d);
  if (!rc) rc = adj_alloc(d);
  if (rc) return rc;
  AdjArrays& A = d->adj->a;
  // the recomputation in the domain's arrays: stage 1 into (V1, bB), the
  // stage-1 face data into the adjoint's own arrays (its reverse needs them
  // after stage 2 has overwritten the domain's), the output into (VN, bN)
  StepTape T;
  for (int q = 0; q < 8; q++) {
    T.V[q] = d->V[q];
    T.V1[q] = d->V1[q];
    T.VN[q] = A.VN[q];
  }
  for (int a = 0; a < 3; a++) {
    T.b[a] = d->bA[a];
    T.b1[a] = d->bB[a];
  }
  for (int dd = 0; dd < 3; dd++) {
    for (int q = 0; q < 5; q++) T.sw[0][dd][q] = A.sw1[dd][q];
    T.sw[1][dd][0] = d->sw.d[dd].SR;
    T.sw[1][dd][1] = d->sw.d[dd].SL;
    for (int c = 0; c < 3; c++) T.sw[1][dd][2 + c] = d->sw.d[dd].dv[c];
  }
  rc = adj_forward(d, T, A.bN);
  if (!rc) rc = adj_reverse(d, T, lam_prim, lam_face);
  return rc;
}

extern "C" int pgpu_adjoint_record(pgpu_domain* d, double* tape)
{
  if (!d || !tape) return fail(PGPU_EINVAL, "NULL argument");
  int rc = adj_checks(d);
  if (rc) return rc;
  const long n = d->g.n;
  const size_t nb = (size_t)n * sizeof(double);
  const StepTape T = tape_view(tape, n);
  hipStream_t s = d->stream;
  for (int q = 0; q < 8; q++)
    if (hipMemcpyAsync(T.V[q], d->V[q], nb, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fail(PGPU_EDEVICE, "tape copy failed");
  for (int a = 0; a < 3; a++)
    if (hipMemcpyAsync(T.b[a], d->bA[a], nb, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fail(PGPU_EDEVICE, "tape copy failed");
  // the step reads its input from the tape, so the output faces can go
  // straight to the domain; the output primitives are copied there after
  rc = adj_forward(d, T, d->bA);
  if (rc) return rc;
  for (int q = 0; q < 8; q++)
    if (hipMemcpyAsync(d->V[q], T.VN[q], nb, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fail(PGPU_EDEVICE, "state copy failed");
  return PGPU_OK;
}

extern "C" int pgpu_adjoint_step_taped(pgpu_domain* d, const double* tape, double* lam_prim,
                                       double* lam_face)
{
  if (!d || !tape || !lam_prim || !lam_face) return fail(PGPU_EINVAL, "NULL argument");
  int rc = adj_checks(d);
  if (!rc) rc = adj_alloc(d);
  if (rc) return rc;
  return adj_reverse(d, tape_view(const_cast<double*>(tape), d->g.n), lam_prim, lam_face);
}

// ---------------------------------------------------------------------
// point-level tangent and adjoint of HLLD (for checking the adjoint)
// ---------------------------------------------------------------------

namespace pgpu {
namespace {

constexpr int kHlldOut = 8;   // rho, mx, my, mz, E fluxes, press, SL, SR

template <int D>
__global__ void k_hlld_tangent(int n, double gamma, const double* vL, const double* vR,
                               const double* tL, const double* tR, double* out, double* dout)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  Dual L[NV], R[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    L[q] = Dual(vL[(long)r * NV + q], tL[(long)r * NV + q]);
    R[q] = Dual(vR[(long)r * NV + q], tR[(long)r * NV + q]);
  }
  RiemannOut<Dual> o;
  hlld<D, false>(L, R, Dual(gamma), Dual(gamma - 1.0), o);
  const Dual res[kHlldOut] = {o.f[RHO], o.f[VX1], o.f[VX2], o.f[VX3], o.f[ENG], o.press, o.SL, o.SR};
#pragma unroll
  for (int q = 0; q < kHlldOut; q++) {
    out[(long)r * kHlldOut + q] = res[q].v;
    dout[(long)r * kHlldOut + q] = res[q].d;
  }
}

template <int D>
__global__ void k_hlld_adjoint(int n, double gamma, const double* vL, const double* vR,
                               const double* lam, double* aL, double* aR)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double L[NV], R[NV], a_L[NV] = {0}, a_R[NV] = {0};
#pragma unroll
  for (int q = 0; q < NV; q++) {
    L[q] = vL[(long)r * NV + q];
    R[q] = vR[(long)r * NV + q];
  }
  HlldSeed s;
  const double* l = lam + (long)r * kHlldOut;
#pragma unroll
  for (int q = 0; q < NV; q++) s.f[q] = 0.0;
  s.f[RHO] = l[0];
  s.f[VX1] = l[1];
  s.f[VX2] = l[2];
  s.f[VX3] = l[3];
  s.f[ENG] = l[4];
  s.press = l[5];
  s.SL = l[6];
  s.SR = l[7];
  hlld_vjp<D>(L, R, gamma, gamma - 1.0, s, a_L, a_R);
#pragma unroll
  for (int q = 0; q < NV; q++) {
    aL[(long)r * NV + q] = a_L[q];
    aR[(long)r * NV + q] = a_R[q];
  }
}

// copies host rows to a temporary device block and back
struct RowScratch {
  void* p = nullptr;
  ~RowScratch() { if (p) (void)hipFree(p); }
};

}  // namespace
}  // namespace pgpu

extern "C" int pgpu_hlld_tangent_rows(int dir, int n, double gamma, const double* vL,
                                      const double* vR, const double* tL, const double* tR,
                                      double* out, double* dout)
{
  if (dir < 0 || dir > 2 || n < 0) return fail(PGPU_EINVAL, "bad direction or count");
  if (n == 0) return PGPU_OK;
  if (!vL || !vR || !tL || !tR || !out || !dout) return fail(PGPU_EINVAL, "NULL argument");
  const size_t in_b = (size_t)n * NV * sizeof(double), out_b = (size_t)n * kHlldOut * sizeof(double);
  RowScratch m;
  if (hipMalloc(&m.p, 4 * in_b + 2 * out_b) != hipSuccess) return fail(PGPU_ENOMEM, "hipMalloc failed");
  double* base = (double*)m.p;
  double *dL = base, *dR = dL + (size_t)n * NV, *dtL = dR + (size_t)n * NV, *dtR = dtL + (size_t)n * NV;
  double *dout_v = dtR + (size_t)n * NV, *dout_d = dout_v + (size_t)n * kHlldOut;
  if (hipMemcpy(dL, vL, in_b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dR, vR, in_b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dtL, tL, in_b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dtR, tR, in_b, hipMemcpyHostToDevice) != hipSuccess)
    return fail(PGPU_EDEVICE, "row upload failed");
  const dim3 b(64), gr((n + 63) / 64);
  if (dir == 0) hipLaunchKernelGGL(k_hlld_tangent<0>, gr, b, 0, 0, n, gamma, dL, dR, dtL, dtR, dout_v, dout_d);
  if (dir == 1) hipLaunchKernelGGL(k_hlld_tangent<1>, gr, b, 0, 0, n, gamma, dL, dR, dtL, dtR, dout_v, dout_d);
  if (dir == 2) hipLaunchKernelGGL(k_hlld_tangent<2>, gr, b, 0, 0, n, gamma, dL, dR, dtL, dtR, dout_v, dout_d);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpy(out, dout_v, out_b, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(dout, dout_d, out_b, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(PGPU_EDEVICE, "tangent kernel failed");
  return PGPU_OK;
}

extern "C" int pgpu_hlld_adjoint_rows(int dir, int n, double gamma, const double* vL,
                                      const double* vR, const double* lam, double* aL,
                                      double* aR)
{
  if (dir < 0 || dir > 2 || n < 0) return fail(PGPU_EINVAL, "bad direction or count");
  if (n == 0) return PGPU_OK;
  if (!vL || !vR || !lam || !aL || !aR) return fail(PGPU_EINVAL, "NULL argument");
  const size_t in_b = (size_t)n * NV * sizeof(double), l_b = (size_t)n * kHlldOut * sizeof(double);
  RowScratch m;
  if (hipMalloc(&m.p, 4 * in_b + l_b) != hipSuccess) return fail(PGPU_ENOMEM, "hipMalloc failed");
  double* base = (double*)m.p;
  double *dL = base, *dR = dL + (size_t)n * NV, *daL = dR + (size_t)n * NV, *daR = daL + (size_t)n * NV;
  double* dlam = daR + (size_t)n * NV;
  if (hipMemcpy(dL, vL, in_b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dR, vR, in_b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dlam, lam, l_b, hipMemcpyHostToDevice) != hipSuccess)
    return fail(PGPU_EDEVICE, "row upload failed");
  const dim3 b(64), gr((n + 63) / 64);
  if (dir == 0) hipLaunchKernelGGL(k_hlld_adjoint<0>, gr, b, 0, 0, n, gamma, dL, dR, dlam, daL, daR);
  if (dir == 1) hipLaunchKernelGGL(k_hlld_adjoint<1>, gr, b, 0, 0, n, gamma, dL, dR, dlam, daL, daR);
  if (dir == 2) hipLaunchKernelGGL(k_hlld_adjoint<2>, gr, b, 0, 0, n, gamma, dL, dR, dlam, daL, daR);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpy(aL, daL, in_b, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(aR, daR, in_b, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(PGPU_EDEVICE, "adjoint kernel failed");
  return PGPU_OK;
}
