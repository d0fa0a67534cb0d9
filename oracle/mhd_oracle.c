/* FP64 CPU restatement of PLUTO's unsplit MHD stage -- the ORACLE.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against; it is never linked into, called by, or used as a
 * fallback for the product library.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py load it.
 *
 * What it restates (PLUTO 4.3, configuration: 3D or 2D ideal MHD,
 * Cartesian uniform grid, PLM with the MC limiter on every variable, HLLD,
 * RK2 unsplit, constrained transport with the UCT_HLL EMF average, periodic
 * boundaries, gamma = 5/3):
 *
 *   ora_rk2_step      AdvanceStep            Time_Stepping/rk_step.c:27-254
 *     stage loop      UpdateStage            Time_Stepping/update_stage.c:37-314
 *     plm_mc          States (MC_LIM)        States/plm_states.c:80-312,
 *                                            States/plm_coeffs.h:119-123
 *     prim2cons       PrimToCons             MHD/mappers.c:25-86
 *     cons2prim       ConsToPrim             MHD/mappers.c:88-254
 *     hlld            HLLD_Solver            MHD/hlld.c:44-452
 *     hll             HLL_Solver             MHD/hll.c:30-169
 *     hllc            HLLC_Solver            MHD/hllc.c:32-230
 *     tvdlf           LF_Solver ("tvdlf")    MHD/tvdlf.c:37-163
 *     roe             Roe_Solver ("roe")     MHD/roe.c:53-700 (HLL_HYBRIDIZATION
 *                                            and CHECK_ROE_MATRIX NO, roe.c:49-50)
 *     flux1           Flux                   MHD/fluxes.c:110-249
 *     signal_speed    MaxSignalSpeed         MHD/eigenv.c:35-104
 *                     HLL_Speed (Davis)      MHD/hll_speed.c:23-108
 *                     SoundSpeed2            EOS/Ideal/eos.c:16-42
 *     (row update)    RightHandSide          MHD/rhs.c:84-201 (USE_PRS_GRADIENT
 *                                            is YES, rhs.c:72-80: the total
 *                                            pressure enters as a separate
 *                                            gradient term, rhs.c:198-199)
 *     (EMF data)      CT_StoreUpwindEMF      MHD/CT/ct_emf.c:104-207
 *                     CT_StoreVelSlopes      MHD/CT/ct_stag_slopes.c:5-49
 *     emf_uct_hll     CT_EMF_HLL_Solver      MHD/CT/ct_emf_average.c:179-352
 *     mc_lim2         CT_GetStagSlopes       MHD/CT/ct_stag_slopes.c:51-106
 *     (face update)   CT_Update              MHD/CT/ct_update.c:29-220
 *     (face average)  CT_AverageMagneticField MHD/CT/ct_field_average.c:13-132
 *   ora_next_dt       NextTimeStep           main.c:389-575
 *
 * Every floating-point expression keeps the reference's operand order so
 * that, built without FMA contraction, results are bit-identical to the
 * reference on x86-64.
 *
 * Layout: periodic box without ghost zones.  prim[8][nz][ny][nx] in the
 * order rho, vx, vy, vz, Bx, By, Bz, p (cell-centred B is always the face
 * average); bx[k][j][i] is the x-face at x_{i+1/2}, by at y_{j+1/2}, bz at
 * z_{k+1/2}.  Periodic ghost cells of the reference are index wraps here:
 * every value PLUTO computes in a ghost zone is, by translation invariance,
 * the value it computes at the periodic image.
 *
 * dims == 2: x-y problem with nz == 1; no z sweep, no Ex/Ey, C_dt divided
 * by 2 (DIMENSIONS).  With vz = Bz = 0 this reproduces the reference's
 * COMPONENTS = 2 build: the extra terms it omits are exact zeros here.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NV 8
enum { RHO = 0, VX1 = 1, VX2 = 2, VX3 = 3, BX1 = 4, BX2 = 5, BX3 = 6, PRS = 7 };
#define ENG PRS
#define FLAG_CONS2PRIM_FAIL 64

static const double small_density  = 1.e-12;   /* globals.h:113 */
static const double small_pressure = 1.e-12;   /* globals.h:114 */

#define MAXD(a, b) ((a) >= (b) ? (a) : (b))    /* macros.h:140 */
#define MIND(a, b) ((a) <= (b) ? (a) : (b))    /* macros.h:143 */
#define ABS_MIN(a, b) (fabs(a) < fabs(b) ? (a) : (b))   /* macros.h:146 */

typedef struct { int vn, vt, vb, bn, bt, bb; } Perm;

/* SetVectorIndices (set_indexes.c:49-123): x: (x,y,z), y: (y,x,z), z: (z,x,y) */
static Perm perm_of (int dir)
{
  static const int t[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};
  Perm p;
  p.vn = VX1 + t[dir][0]; p.vt = VX1 + t[dir][1]; p.vb = VX1 + t[dir][2];
  p.bn = BX1 + t[dir][0]; p.bt = BX1 + t[dir][1]; p.bb = BX1 + t[dir][2];
  return p;
}

/* ------------------------------------------------------------------ */
/* Conversions                                                          */
/* ------------------------------------------------------------------ */

static void prim2cons (const double *v, double *u, double gmm1)
{
  double kinb2;
  u[RHO] = v[RHO];
  u[VX1] = v[RHO]*v[VX1];
  u[VX2] = v[RHO]*v[VX2];
  u[VX3] = v[RHO]*v[VX3];
  u[BX1] = v[BX1];
  u[BX2] = v[BX2];
  u[BX3] = v[BX3];
  kinb2 = v[VX1]*v[VX1] + v[VX2]*v[VX2] + v[VX3]*v[VX3];
  kinb2 = v[RHO]*kinb2 + v[BX1]*v[BX1] + v[BX2]*v[BX2] + v[BX3]*v[BX3];
  kinb2 *= 0.5;
  u[ENG] = kinb2 + v[PRS]/gmm1;
}

/* Returns 1 on a floor; u may be modified (floors), as in the reference. */
static int cons2prim (double *u, double *v, double gmm1, unsigned char *flag)
{
  int fail = 0;
  double m2, b2, tau, kinb2;
  m2 = u[VX1]*u[VX1] + u[VX2]*u[VX2] + u[VX3]*u[VX3];
  b2 = u[BX1]*u[BX1] + u[BX2]*u[BX2] + u[BX3]*u[BX3];
  if (u[RHO] < 0.0){
    u[RHO] = small_density;
    *flag |= FLAG_CONS2PRIM_FAIL;
    fail = 1;
  }
  v[RHO] = u[RHO];
  tau = 1.0/u[RHO];
  v[VX1] = u[VX1]*tau;
  v[VX2] = u[VX2]*tau;
  v[VX3] = u[VX3]*tau;
  v[BX1] = u[BX1];
  v[BX2] = u[BX2];
  v[BX3] = u[BX3];
  kinb2 = 0.5*(m2*tau + b2);
  if (u[ENG] < 0.0){
    u[ENG] = small_pressure/gmm1 + kinb2;
    *flag |= FLAG_CONS2PRIM_FAIL;
    fail = 1;
  }
  v[PRS] = gmm1*(u[ENG] - kinb2);
  if (v[PRS] < 0.0){
    v[PRS] = small_pressure;
    u[ENG] = v[PRS]/gmm1 + kinb2;
    *flag |= FLAG_CONS2PRIM_FAIL;
    fail = 1;
  }
  return fail;
}

/* ------------------------------------------------------------------ */
/* PLM with the MC limiter on a uniform Cartesian grid                  */
/* ------------------------------------------------------------------ */

/* vm1, v0, vp1: primitive vectors of cells i-1, i, i+1.
   Returns vp = state at i+1/2 and vm = state at i-1/2. */
static void plm_mc (const double *vm1, const double *v0, const double *vp1,
                    double *vp, double *vm)
{
  int nv;
  for (nv = 0; nv < NV; nv++){
    double dvp = vp1[nv] - v0[nv];
    double dvm = v0[nv] - vm1[nv];
    double dvl;
    if (dvp*dvm > 0.0){
      double qc = 0.5*(dvm + dvp), scrh = 2.0*ABS_MIN(dvp, dvm);
      dvl = ABS_MIN(qc, scrh);
    } else dvl = 0.0;
    vp[nv] = v0[nv] + dvl*0.5;
    vm[nv] = v0[nv] - dvl*0.5;
  }
}

/* ------------------------------------------------------------------ */
/* Physical flux, fast speed, HLLD                                      */
/* ------------------------------------------------------------------ */

static void flux1 (const double *v, const double *u, Perm p, double *fx, double *ptot)
{
  double bmag2, vb;
  bmag2 = v[BX1]*v[BX1] + v[BX2]*v[BX2] + v[BX3]*v[BX3];
  *ptot = v[PRS] + 0.5*bmag2;
  vb = v[VX1]*v[BX1] + v[VX2]*v[BX2] + v[VX3]*v[BX3];
  fx[RHO] = u[p.vn];
  fx[VX1] = v[p.vn]*u[VX1] - v[p.bn]*v[BX1];
  fx[VX2] = v[p.vn]*u[VX2] - v[p.bn]*v[BX2];
  fx[VX3] = v[p.vn]*u[VX3] - v[p.bn]*v[BX3];
  fx[p.bn] = 0.0;
  fx[p.bt] = v[p.vn]*v[p.bt] - v[p.bn]*v[p.vt];
  fx[p.bb] = v[p.vn]*v[p.bb] - v[p.bn]*v[p.vb];
  fx[ENG] = (u[ENG] + *ptot)*v[p.vn] - v[p.bn]*vb;
}

static void signal_speed (const double *v, Perm p, double gamma,
                          double *cmin, double *cmax)
{
  double gpr, b1, b2, b3, btmag2, bmag2, cf;
  gpr = gamma*v[PRS];
  b1 = v[p.bn]; b2 = v[p.bt]; b3 = v[p.bb];
  btmag2 = b2*b2 + b3*b3;
  bmag2  = b1*b1 + btmag2;
  cf = gpr - bmag2;
  cf = gpr + bmag2 + sqrt(cf*cf + 4.0*gpr*btmag2);
  cf = sqrt(0.5*cf/v[RHO]);
  *cmin = v[p.vn] - cf;
  *cmax = v[p.vn] + cf;
}

typedef struct {
  double flux[NV], press, SL, SR, cmax, mach;
} HLLDOut;

static void hlld (const double *vL, const double *vR, Perm p, double gamma,
                  HLLDOut *o)
{
  int nv, revert_to_hllc;
  double gmm1 = gamma - 1.0;
  double uL[NV], uR[NV], fL[NV], fR[NV], ptL, ptR, a2L, a2R;
  double slmin, slmax, srmin, srmax, SL, SR, scrh;
  double usL[NV], usR[NV], ussl[NV], ussr[NV], Uhll[NV];
  double vsL, wsL, scrhL, S1L, sqrL, duL;
  double vsR, wsR, scrhR, S1R, sqrR, duR;
  double Bx, Bx1, SM, sBx, pts, vss, wss;

  prim2cons (vL, uL, gmm1);
  prim2cons (vR, uR, gmm1);
  a2L = gamma*vL[PRS]/vL[RHO];
  a2R = gamma*vR[PRS]/vR[RHO];
  flux1 (vL, uL, p, fL, &ptL);
  flux1 (vR, uR, p, fR, &ptR);

  signal_speed (vL, p, gamma, &slmin, &slmax);
  signal_speed (vR, p, gamma, &srmin, &srmax);
  SL = MIND(slmin, srmin);
  SR = MAXD(slmax, srmax);
  scrh  = fabs(vL[p.vn]) + fabs(vR[p.vn]);
  scrh /= sqrt(a2L) + sqrt(a2R);
  o->mach = scrh;
  o->SL = SL; o->SR = SR;
  o->cmax = MAXD(fabs(SL), fabs(SR));

  if (SL >= 0.0){
    for (nv = 0; nv < NV; nv++) o->flux[nv] = fL[nv];
    o->press = ptL;
    return;
  }
  if (SR <= 0.0){
    for (nv = 0; nv < NV; nv++) o->flux[nv] = fR[nv];
    o->press = ptR;
    return;
  }

  scrh = 1.0/(SR - SL);
  Bx1 = Bx = (SR*vR[p.bn] - SL*vL[p.bn])*scrh;
  sBx = (Bx > 0.0 ? 1.0 : -1.0);

  duL = SL - vL[p.vn];
  duR = SR - vR[p.vn];

  scrh = 1.0/(duR*uR[RHO] - duL*uL[RHO]);
  SM   = (duR*uR[p.vn] - duL*uL[p.vn] - ptR + ptL)*scrh;

  pts  = duR*uR[RHO]*ptL - duL*uL[RHO]*ptR +
         vL[RHO]*vR[RHO]*duR*duL*(vR[p.vn]- vL[p.vn]);
  pts *= scrh;

  usL[RHO] = uL[RHO]*duL/(SL - SM);
  usR[RHO] = uR[RHO]*duR/(SR - SM);

  sqrL = sqrt(usL[RHO]);
  sqrR = sqrt(usR[RHO]);

  S1L = SM - fabs(Bx)/sqrL;
  S1R = SM + fabs(Bx)/sqrR;

  revert_to_hllc = 0;
  if ( (S1L - SL) <  1.e-4*(SM - SL) ) revert_to_hllc = 1;
  if ( (S1R - SR) > -1.e-4*(SR - SM) ) revert_to_hllc = 1;

  if (revert_to_hllc){
    scrh = 1.0/(SR - SL);
    for (nv = 0; nv < NV; nv++){
      Uhll[nv]  = SR*uR[nv] - SL*uL[nv] + fL[nv] - fR[nv];
      Uhll[nv] *= scrh;
    }
    usL[p.bn] = usR[p.bn] = Uhll[p.bn];
    usL[p.bt] = usR[p.bt] = Uhll[p.bt];
    usL[p.bb] = usR[p.bb] = Uhll[p.bb];
    S1L = S1R = SM;
  }else{
    scrhL = (uL[RHO]*duL*duL - Bx*Bx)/(uL[RHO]*duL*(SL - SM) - Bx*Bx);
    scrhR = (uR[RHO]*duR*duR - Bx*Bx)/(uR[RHO]*duR*(SR - SM) - Bx*Bx);
    usL[p.bn] = Bx1;
    usL[p.bt] = uL[p.bt]*scrhL;
    usL[p.bb] = uL[p.bb]*scrhL;
    usR[p.bn] = Bx1;
    usR[p.bt] = uR[p.bt]*scrhR;
    usR[p.bb] = uR[p.bb]*scrhR;
  }

  scrhL = Bx/(uL[RHO]*duL);
  scrhR = Bx/(uR[RHO]*duR);

  vsL = vL[p.vt] - scrhL*(usL[p.bt] - uL[p.bt]);
  vsR = vR[p.vt] - scrhR*(usR[p.bt] - uR[p.bt]);
  wsL = vL[p.vb] - scrhL*(usL[p.bb] - uL[p.bb]);
  wsR = vR[p.vb] - scrhR*(usR[p.bb] - uR[p.bb]);

  usL[p.vn] = usL[RHO]*SM;
  usR[p.vn] = usR[RHO]*SM;
  usL[p.vt] = usL[RHO]*vsL;
  usR[p.vt] = usR[RHO]*vsR;
  usL[p.vb] = usL[RHO]*wsL;
  usR[p.vb] = usR[RHO]*wsR;

  scrhL  = vL[p.vn]*Bx1 + vL[p.vt]*uL[p.bt] + vL[p.vb]*uL[p.bb];
  scrhL -= SM*Bx1 + vsL*usL[p.bt] + wsL*usL[p.bb];
  usL[ENG]  = duL*uL[ENG] - ptL*vL[p.vn] + pts*SM + Bx*scrhL;
  usL[ENG] /= SL - SM;

  scrhR  = vR[p.vn]*Bx1 + vR[p.vt]*uR[p.bt] + vR[p.vb]*uR[p.bb];
  scrhR -= SM*Bx1 + vsR*usR[p.bt] + wsR*usR[p.bb];
  usR[ENG]  = duR*uR[ENG] - ptR*vR[p.vn] + pts*SM + Bx*scrhR;
  usR[ENG] /= SR - SM;

  if (S1L >= 0.0){
    for (nv = 0; nv < NV; nv++) o->flux[nv] = fL[nv] + SL*(usL[nv] - uL[nv]);
    o->press = ptL;
  }else if (S1R <= 0.0){
    for (nv = 0; nv < NV; nv++) o->flux[nv] = fR[nv] + SR*(usR[nv] - uR[nv]);
    o->press = ptR;
  }else{
    ussl[RHO] = usL[RHO];
    ussr[RHO] = usR[RHO];

    vss  = sqrL*vsL + sqrR*vsR + (usR[p.bt] - usL[p.bt])*sBx;
    vss /= sqrL + sqrR;
    wss  = sqrL*wsL + sqrR*wsR + (usR[p.bb] - usL[p.bb])*sBx;
    wss /= sqrL + sqrR;

    ussl[p.vn] = ussl[RHO]*SM;
    ussr[p.vn] = ussr[RHO]*SM;
    ussl[p.vt] = ussl[RHO]*vss;
    ussr[p.vt] = ussr[RHO]*vss;
    ussl[p.vb] = ussl[RHO]*wss;
    ussr[p.vb] = ussr[RHO]*wss;

    ussl[p.bn] = ussr[p.bn] = Bx1;
    ussl[p.bt]  = sqrL*usR[p.bt] + sqrR*usL[p.bt] + sqrL*sqrR*(vsR - vsL)*sBx;
    ussl[p.bt] /= sqrL + sqrR;
    ussr[p.bt]  = ussl[p.bt];
    ussl[p.bb]  = sqrL*usR[p.bb] + sqrR*usL[p.bb] + sqrL*sqrR*(wsR - wsL)*sBx;
    ussl[p.bb] /= sqrL + sqrR;
    ussr[p.bb]  = ussl[p.bb];

    scrhL  = SM*Bx1 + vsL*usL [p.bt] + wsL*usL [p.bb];
    scrhL -= SM*Bx1 + vss*ussl[p.bt] + wss*ussl[p.bb];
    scrhR  = SM*Bx1 + vsR*usR [p.bt] + wsR*usR [p.bb];
    scrhR -= SM*Bx1 + vss*ussr[p.bt] + wss*ussr[p.bb];

    ussl[ENG] = usL[ENG] - sqrL*scrhL*sBx;
    ussr[ENG] = usR[ENG] + sqrR*scrhR*sBx;

    if (SM >= 0.0){
      for (nv = 0; nv < NV; nv++)
        o->flux[nv] = fL[nv] + S1L*(ussl[nv] - usL[nv]) + SL*(usL[nv] - uL[nv]);
      o->press = ptL;
    }else{
      for (nv = 0; nv < NV; nv++)
        o->flux[nv] = fR[nv] + S1R*(ussr[nv] - usR[nv]) + SR*(usR[nv] - uR[nv]);
      o->press = ptR;
    }
  }
}

/* HLL_Solver (MHD/hll.c:30-169): the same Flux, SoundSpeed2 and HLL_Speed
   (with its Mach estimate, hll_speed.c:85-105) as hlld above; the outer
   states for SL > 0 and SR < 0 (strict, hll.c:117,123), otherwise the
   single intermediate state's flux (hll.c:131-137). */
static void hll (const double *vL, const double *vR, Perm p, double gamma,
                 HLLDOut *o)
{
  int nv;
  double gmm1 = gamma - 1.0;
  double uL[NV], uR[NV], fL[NV], fR[NV], ptL, ptR, a2L, a2R;
  double slmin, slmax, srmin, srmax, SL, SR, scrh;

  prim2cons (vL, uL, gmm1);
  prim2cons (vR, uR, gmm1);
  a2L = gamma*vL[PRS]/vL[RHO];
  a2R = gamma*vR[PRS]/vR[RHO];
  flux1 (vL, uL, p, fL, &ptL);
  flux1 (vR, uR, p, fR, &ptR);

  signal_speed (vL, p, gamma, &slmin, &slmax);
  signal_speed (vR, p, gamma, &srmin, &srmax);
  SL = MIND(slmin, srmin);
  SR = MAXD(slmax, srmax);
  scrh  = fabs(vL[p.vn]) + fabs(vR[p.vn]);
  scrh /= sqrt(a2L) + sqrt(a2R);
  o->mach = scrh;
  o->SL = SL; o->SR = SR;
  o->cmax = MAXD(fabs(SL), fabs(SR));

  if (SL > 0.0){
    for (nv = 0; nv < NV; nv++) o->flux[nv] = fL[nv];
    o->press = ptL;
  }else if (SR < 0.0){
    for (nv = 0; nv < NV; nv++) o->flux[nv] = fR[nv];
    o->press = ptR;
  }else{
    scrh = 1.0/(SR - SL);
    for (nv = 0; nv < NV; nv++){
      o->flux[nv]  = SL*SR*(uR[nv] - uL[nv]) + SR*fL[nv] - SL*fR[nv];
      o->flux[nv] *= scrh;
    }
    o->press = (SR*ptL - SL*ptR)*scrh;
  }
}

/* HLLC_Solver (MHD/hllc.c:32-230), Li (2005) as PLUTO modifies it: same
   Flux, SoundSpeed2 and HLL_Speed as hlld; outer states for SL >= 0 and
   SR <= 0; else the HLL state and flux give vxs, ps and the star field, and
   the star state on the contact's side gives the flux. */
static void hllc (const double *vL, const double *vR, Perm p, double gamma,
                  HLLDOut *o)
{
  int nv;
  double gmm1 = gamma - 1.0;
  double uL[NV], uR[NV], fL[NV], fR[NV], ptL, ptR, a2L, a2R;
  double slmin, slmax, srmin, srmax, SL, SR, scrh;
  double Uhll[NV], Fhll[NV], usl[NV], usr[NV];
  double pl, pr, vBl, vBr, vxl, vxr, Bxs, Bys, Bzs, vxs, ps, vBs;

  prim2cons (vL, uL, gmm1);
  prim2cons (vR, uR, gmm1);
  a2L = gamma*vL[PRS]/vL[RHO];
  a2R = gamma*vR[PRS]/vR[RHO];
  flux1 (vL, uL, p, fL, &ptL);
  flux1 (vR, uR, p, fR, &ptR);

  signal_speed (vL, p, gamma, &slmin, &slmax);
  signal_speed (vR, p, gamma, &srmin, &srmax);
  SL = MIND(slmin, srmin);
  SR = MAXD(slmax, srmax);
  scrh  = fabs(vL[p.vn]) + fabs(vR[p.vn]);
  scrh /= sqrt(a2L) + sqrt(a2R);
  o->mach = scrh;
  o->SL = SL; o->SR = SR;
  o->cmax = MAXD(fabs(SL), fabs(SR));

  if (SL >= 0.0){
    for (nv = 0; nv < NV; nv++) o->flux[nv] = fL[nv];
    o->press = ptL;
    return;
  }
  if (SR <= 0.0){
    for (nv = 0; nv < NV; nv++) o->flux[nv] = fR[nv];
    o->press = ptR;
    return;
  }
  scrh = 1.0/(SR - SL);
  for (nv = 0; nv < NV; nv++){
    Uhll[nv]  = SR*uR[nv] - SL*uL[nv] + fL[nv] - fR[nv];
    Uhll[nv] *= scrh;
    Fhll[nv]  = SL*SR*(uR[nv] - uL[nv]) + SR*fL[nv] - SL*fR[nv];
    Fhll[nv] *= scrh;
  }
  Uhll[p.vn] += (ptL - ptR)*scrh;
  Fhll[p.vn] += (SR*ptL - SL*ptR)*scrh;

  pl = vL[BX1]*vL[BX1] + vL[BX2]*vL[BX2] + vL[BX3]*vL[BX3];
  pr = vR[BX1]*vR[BX1] + vR[BX2]*vR[BX2] + vR[BX3]*vR[BX3];
  pl = vL[PRS] + 0.5*pl;
  pr = vR[PRS] + 0.5*pr;
  vBl = vL[VX1]*vL[BX1] + vL[VX2]*vL[BX2] + vL[VX3]*vL[BX3];
  vBr = vR[VX1]*vR[BX1] + vR[VX2]*vR[BX2] + vR[VX3]*vR[BX3];
  vxl = vL[p.vn];
  vxr = vR[p.vn];
  Bxs = Uhll[p.bn]; Bys = Uhll[p.bt]; Bzs = Uhll[p.bb];
  vxs = Uhll[p.vn]/Uhll[RHO];
  ps  = Fhll[p.vn] + Bxs*Bxs - Fhll[RHO]*vxs;
  vBs = Uhll[BX1]*Uhll[VX1] + Uhll[BX2]*Uhll[VX2] + Uhll[BX3]*Uhll[VX3];
  vBs /= Uhll[RHO];

  usl[RHO] = uL[RHO]*(SL - vxl)/(SL - vxs);
  usr[RHO] = uR[RHO]*(SR - vxr)/(SR - vxs);
  usl[ENG] = (uL[ENG]*(SL - vxl) + ps*vxs - pl*vxl - Bxs*vBs + vL[p.bn]*vBl)/(SL - vxs);
  usr[ENG] = (uR[ENG]*(SR - vxr) + ps*vxs - pr*vxr - Bxs*vBs + vR[p.bn]*vBr)/(SR - vxs);
  usl[p.vn] = usl[RHO]*vxs;
  usr[p.vn] = usr[RHO]*vxs;
  usl[p.vt] = (uL[p.vt]*(SL - vxl) - (Bxs*Bys - vL[p.bn]*vL[p.bt]))/(SL - vxs);
  usr[p.vt] = (uR[p.vt]*(SR - vxr) - (Bxs*Bys - vR[p.bn]*vR[p.bt]))/(SR - vxs);
  usl[p.vb] = (uL[p.vb]*(SL - vxl) - (Bxs*Bzs - vL[p.bn]*vL[p.bb]))/(SL - vxs);
  usr[p.vb] = (uR[p.vb]*(SR - vxr) - (Bxs*Bzs - vR[p.bn]*vR[p.bb]))/(SR - vxs);
  usl[p.bn] = usr[p.bn] = Bxs;
  usl[p.bt] = usr[p.bt] = Bys;
  usl[p.bb] = usr[p.bb] = Bzs;

  if (vxs >= 0.0){
    for (nv = 0; nv < NV; nv++) o->flux[nv] = fL[nv] + SL*(usl[nv] - uL[nv]);
    o->press = ptL;
  }else{
    for (nv = 0; nv < NV; nv++) o->flux[nv] = fR[nv] + SR*(usr[nv] - uR[nv]);
    o->press = ptR;
  }
}

/* Riemann solver of the sweeps: 0 = HLLD, 1 = HLL, 2 = HLLC (set per call
   from Python through ora_set_solver). */
/* LF_Solver ("tvdlf", MHD/tvdlf.c:37-163): the Rusanov flux with the
   largest fast speed |v_n +- c_f| of the arithmetic mean of the two states
   (MaxSignalSpeed, eigenv.c:35-104); the Mach estimate is that of the mean
   state. */
static void tvdlf (const double *vL, const double *vR, Perm p, double gamma,
                   HLLDOut *o)
{
  int nv;
  double gmm1 = gamma - 1.0;
  double uL[NV], uR[NV], fL[NV], fR[NV], ptL, ptR, vRL[NV], cmin, cmax, cRL;

  prim2cons (vL, uL, gmm1);
  prim2cons (vR, uR, gmm1);
  flux1 (vL, uL, p, fL, &ptL);
  flux1 (vR, uR, p, fR, &ptR);
  for (nv = 0; nv < NV; nv++) vRL[nv] = 0.5*(vL[nv] + vR[nv]);
  o->mach = fabs(vRL[p.vn])/sqrt(gamma*vRL[PRS]/vRL[RHO]);
  signal_speed (vRL, p, gamma, &cmin, &cmax);
  cRL = MAXD(fabs(cmin), fabs(cmax));
  o->SL = -cRL;
  o->SR =  cRL;
  o->cmax = cRL;
  for (nv = 0; nv < NV; nv++) o->flux[nv] = 0.5*(fL[nv] + fR[nv] - cRL*(uR[nv] - uL[nv]));
  o->press = 0.5*(ptL + ptR);
}

/* Roe_Solver ("roe", MHD/roe.c:53-700): Cargo & Gallice (1997) averages,
   the conservative eigenvectors and wave strengths of roe.c:383-604 in the
   wave order of mod_defs.h:54-68 (fast-, fast+, entropy, div.B, slow-,
   slow+, Alfven-, Alfven+), the entropy fix of roe.c:616-631 and the flux
   sum of roe.c:636-642.  two_comp: the reference's COMPONENTS 2 build (2D
   boxes): beta_y = DSIGN(By) (1 when Btmag <= 1e-9) and no Alfven waves. */
static int g_two_comp = 0;

static void roe (const double *vL, const double *vR, Perm p, double gamma,
                 HLLDOut *o)
{
  enum { KFM = 0, KFP = 1, KEN = 2, KDB = 3, KSM = 4, KSP = 5, KAM = 6, KAP = 7 };
  const int MXn = p.vn, MXt = p.vt, MXb = p.vb, VXn = p.vn, VXt = p.vt, VXb = p.vb;
  const int BXn = p.bn, BXt = p.bt, BXb = p.bb;
  const int nwave = g_two_comp ? 6 : 8;
  const double delta = 1.e-6, g1 = gamma - 1.0, sqrt_1_2 = 0.70710678118654752440;
  double uL[NV], uR[NV], fL[NV], fR[NV], pL, pR;
  double dV[NV], dU[NV], Rc[NV][NV], eta[NV], lambda[NV], alambda[NV];
  double sqr_rho_L, sqr_rho_R, sl, sr, rho, sqrt_rho, u, v, w, Bx, By, Bz, sBx;
  double bx, by, bz, bt2, b2, Btmag, X, vdm, BdB, vel2, HL, HR, H, Hgas, a2;
  double scrh, ca2, cf2, cs2, cf, cs, ca, a, alpha_f, alpha_s, beta_y, beta_z;
  double beta_dv, beta_dB, beta_v;
  int nv, k;

  prim2cons (vL, uL, g1);
  prim2cons (vR, uR, g1);
  flux1 (vL, uL, p, fL, &pL);
  flux1 (vR, uR, p, fR, &pR);
  memset (Rc, 0, sizeof(Rc));
  for (nv = 0; nv < NV; nv++){
    dV[nv] = vR[nv] - vL[nv];
    dU[nv] = uR[nv] - uL[nv];
  }
  sqr_rho_L = sqrt(vL[RHO]);
  sqr_rho_R = sqrt(vR[RHO]);
  sl = sqr_rho_L/(sqr_rho_L + sqr_rho_R);
  sr = sqr_rho_R/(sqr_rho_L + sqr_rho_R);
  rho = sr*vL[RHO] + sl*vR[RHO];
  sqrt_rho = sqrt(rho);
  u = sl*vL[VXn] + sr*vR[VXn];
  v = sl*vL[VXt] + sr*vR[VXt];
  w = sl*vL[VXb] + sr*vR[VXb];
  Bx = sr*vL[BXn] + sl*vR[BXn];
  By = sr*vL[BXt] + sl*vR[BXt];
  Bz = sr*vL[BXb] + sl*vR[BXb];
  sBx = (Bx >= 0.0 ? 1.0 : -1.0);
  bx = Bx/sqrt_rho;
  by = By/sqrt_rho;
  bz = Bz/sqrt_rho;
  bt2   = 0.0 + by*by + bz*bz;
  b2    = bx*bx + bt2;
  Btmag = sqrt(bt2*rho);
  X  = dV[BXn]*dV[BXn] + dV[BXt]*dV[BXt] + dV[BXb]*dV[BXb];
  X /= (sqr_rho_L + sqr_rho_R)*(sqr_rho_L + sqr_rho_R)*2.0;
  vdm = u*dU[MXn] + v*dU[MXt] + w*dU[MXb];
  BdB = Bx*dU[BXn] + By*dU[BXt] + Bz*dU[BXb];
  vel2    = u*u + v*v + w*w;
  dV[PRS] = g1*((0.5*vel2 - X)*dV[RHO] - vdm + dU[ENG] - BdB);
  HL   = (uL[ENG] + pL)/vL[RHO];
  HR   = (uR[ENG] + pR)/vR[RHO];
  H    = sl*HL + sr*HR;
  Hgas = H - b2;
  a2 = (2.0 - gamma)*X + g1*(Hgas - 0.5*vel2);
  /* a2 < 0 stops the reference (roe.c:301-308); not reached by the tests */

  scrh = a2 - b2;
  ca2  = bx*bx;
  scrh = scrh*scrh + 4.0*bt2*a2;
  scrh = sqrt(scrh);
  cf2 = 0.5*(a2 + b2 + scrh);
  cs2 = a2*ca2/cf2;
  cf = sqrt(cf2);
  cs = sqrt(cs2);
  ca = sqrt(ca2);
  a  = sqrt(a2);
  if (cf == cs) {
    alpha_f = 1.0; alpha_s = 0.0;
  }else if (a <= cs) {
    alpha_f = 0.0; alpha_s = 1.0;
  }else if (cf <= a){
    alpha_f = 1.0; alpha_s = 0.0;
  }else{
    scrh    = 1.0/(cf2 - cs2);
    alpha_f = (a2  - cs2)*scrh;
    alpha_s = (cf2 -  a2)*scrh;
    alpha_f = MAXD(0.0, alpha_f);
    alpha_s = MAXD(0.0, alpha_s);
    alpha_f = sqrt(alpha_f);
    alpha_s = sqrt(alpha_s);
  }
  if (g_two_comp){
    beta_y = Btmag > 1.e-9 ? (By >= 0.0 ? 1.0 : -1.0) : 1.0;
    beta_z = 0.0;
  }else if (Btmag > 1.e-9){
    beta_y = By/Btmag;
    beta_z = Bz/Btmag;
  }else{
    beta_z = beta_y = sqrt_1_2;
  }

  /* fast wave u - c_f */
  k = KFM;
  lambda[k] = u - cf;
  scrh    = alpha_s*cs*sBx;
  beta_dv = 0.0 + beta_y*dV[VXt] + beta_z*dV[VXb];
  beta_dB = 0.0 + beta_y*dV[BXt] + beta_z*dV[BXb];
  beta_v  = 0.0 + beta_y*v       + beta_z*w;
  Rc[RHO][k] = alpha_f;
  Rc[MXn][k] = alpha_f*lambda[k];
  Rc[MXt][k] = alpha_f*v + scrh*beta_y;
  Rc[MXb][k] = alpha_f*w + scrh*beta_z;
  Rc[BXt][k] = alpha_s*a*beta_y/sqrt_rho;
  Rc[BXb][k] = alpha_s*a*beta_z/sqrt_rho;
  Rc[ENG][k] =   alpha_f*(Hgas - u*cf) + scrh*beta_v
               + alpha_s*a*Btmag/sqrt_rho;
  eta[k] =   alpha_f*(X*dV[RHO] + dV[PRS]) + rho*scrh*beta_dv
           - rho*alpha_f*cf*dV[VXn]        + sqrt_rho*alpha_s*a*beta_dB;
  eta[k] *= 0.5/a2;

  /* fast wave u + c_f */
  k = KFP;
  lambda[k] = u + cf;
  Rc[RHO][k] = alpha_f;
  Rc[MXn][k] = alpha_f*lambda[k];
  Rc[MXt][k] = alpha_f*v - scrh*beta_y;
  Rc[MXb][k] = alpha_f*w - scrh*beta_z;
  Rc[BXt][k] = Rc[BXt][KFM];
  Rc[BXb][k] = Rc[BXb][KFM];
  Rc[ENG][k] =   alpha_f*(Hgas + u*cf) - scrh*beta_v
               + alpha_s*a*Btmag/sqrt_rho;
  eta[k] =   alpha_f*(X*dV[RHO] + dV[PRS]) - rho*scrh*beta_dv
           + rho*alpha_f*cf*dV[VXn]        + sqrt_rho*alpha_s*a*beta_dB;
  eta[k] *= 0.5/a2;

  /* entropy wave u */
  k = KEN;
  lambda[k] = u;
  Rc[RHO][k] = 1.0;
  Rc[MXn][k] = u;
  Rc[MXt][k] = v;
  Rc[MXb][k] = w;
  Rc[ENG][k] = 0.5*vel2 + (gamma - 2.0)/g1*X;
  eta[k] = ((a2 - X)*dV[RHO] - dV[PRS])/a2;

  /* div.B wave u (CT: zero jump) */
  k = KDB;
  lambda[k] = u;
  Rc[BXn][k] = eta[k] = 0.0;

  /* slow wave u - c_s */
  scrh = alpha_f*cf*sBx;
  k = KSM;
  lambda[k] = u - cs;
  Rc[RHO][k] = alpha_s;
  Rc[MXn][k] = alpha_s*lambda[k];
  Rc[MXt][k] = alpha_s*v - scrh*beta_y;
  Rc[MXb][k] = alpha_s*w - scrh*beta_z;
  Rc[BXt][k] = - alpha_f*a*beta_y/sqrt_rho;
  Rc[BXb][k] = - alpha_f*a*beta_z/sqrt_rho;
  Rc[ENG][k] =   alpha_s*(Hgas - u*cs) - scrh*beta_v
               - alpha_f*a*Btmag/sqrt_rho;
  eta[k] =   alpha_s*(X*dV[RHO] + dV[PRS]) - rho*scrh*beta_dv
           - rho*alpha_s*cs*dV[VXn]        - sqrt_rho*alpha_f*a*beta_dB;
  eta[k] *= 0.5/a2;

  /* slow wave u + c_s */
  k = KSP;
  lambda[k] = u + cs;
  Rc[RHO][k] = alpha_s;
  Rc[MXn][k] = alpha_s*lambda[k];
  Rc[MXt][k] = alpha_s*v + scrh*beta_y;
  Rc[MXb][k] = alpha_s*w + scrh*beta_z;
  Rc[BXt][k] = Rc[BXt][KSM];
  Rc[BXb][k] = Rc[BXb][KSM];
  Rc[ENG][k] =   alpha_s*(Hgas + u*cs) + scrh*beta_v
               - alpha_f*a*Btmag/sqrt_rho;
  eta[k] =   alpha_s*(X*dV[RHO] + dV[PRS]) + rho*scrh*beta_dv
           + rho*alpha_s*cs*dV[VXn]        - sqrt_rho*alpha_f*a*beta_dB;
  eta[k] *= 0.5/a2;

  if (!g_two_comp){
    /* Alfven wave u - c_a */
    k = KAM;
    lambda[k] = u - ca;
    Rc[MXt][k] = - rho*beta_z;
    Rc[MXb][k] = + rho*beta_y;
    Rc[BXt][k] = - sBx*sqrt_rho*beta_z;
    Rc[BXb][k] =   sBx*sqrt_rho*beta_y;
    Rc[ENG][k] = - rho*(v*beta_z - w*beta_y);
    eta[k] = + beta_y*dV[VXb]               - beta_z*dV[VXt]
             + sBx/sqrt_rho*(beta_y*dV[BXb] - beta_z*dV[BXt]);
    eta[k] *= 0.5;

    /* Alfven wave u + c_a */
    k = KAP;
    lambda[k] = u + ca;
    Rc[MXt][k] = - Rc[MXt][KAM];
    Rc[MXb][k] = - Rc[MXb][KAM];
    Rc[BXt][k] =   Rc[BXt][KAM];
    Rc[BXb][k] =   Rc[BXb][KAM];
    Rc[ENG][k] = - Rc[ENG][KAM];
    eta[k] = - beta_y*dV[VXb]               + beta_z*dV[VXt]
             + sBx/sqrt_rho*(beta_y*dV[BXb] - beta_z*dV[BXt]);
    eta[k] *= 0.5;
  }

  o->cmax = fabs(u) + cf;
  o->mach = fabs(u/a);
  for (k = 0; k < nwave; k++) alambda[k] = fabs(lambda[k]);
  if (alambda[KFM] < 0.5*delta) alambda[KFM] = lambda[KFM]*lambda[KFM]/delta + 0.25*delta;
  if (alambda[KFP] < 0.5*delta) alambda[KFP] = lambda[KFP]*lambda[KFP]/delta + 0.25*delta;
  if (alambda[KSM] < 0.5*delta) alambda[KSM] = lambda[KSM]*lambda[KSM]/delta + 0.25*delta;
  if (alambda[KSP] < 0.5*delta) alambda[KSP] = lambda[KSP]*lambda[KSP]/delta + 0.25*delta;
  for (nv = 0; nv < NV; nv++){
    scrh = 0.0;
    for (k = 0; k < nwave; k++) scrh += alambda[k]*eta[k]*Rc[nv][k];
    o->flux[nv] = 0.5*(fL[nv] + fR[nv] - scrh);
  }
  o->press = 0.5*(pL + pR);
  o->SL = lambda[KFM];
  o->SR = lambda[KFP];
}

static int g_solver = 0;
void ora_set_solver (int s) { g_solver = s; }

static void riemann (const double *vL, const double *vR, Perm p, double gamma,
                     HLLDOut *o)
{
  if (g_solver == 1) hll (vL, vR, p, gamma, o);
  else if (g_solver == 2) hllc (vL, vR, p, gamma, o);
  else if (g_solver == 3) tvdlf (vL, vR, p, gamma, o);
  else if (g_solver == 4) roe (vL, vR, p, gamma, o);
  else hlld (vL, vR, p, gamma, o);
}

/* ------------------------------------------------------------------ */
/* Grid bookkeeping                                                     */
/* ------------------------------------------------------------------ */

typedef struct {
  int nx, ny, nz, dims;
  long n;
  double dx[3], gamma;
} Box;

static inline long IDX (const Box *g, int i, int j, int k)
{
  i = (i + g->nx) % g->nx;    /* indices stay within one period of the box */
  j = (j + g->ny) % g->ny;
  k = (k + g->nz) % g->nz;
  return ((long)k*g->ny + j)*g->nx + i;
}

/* Per-stage scratch (allocated per call). */
typedef struct {
  double *U;        /* [8][n] conservative, accumulated                 */
  double *S[3][2];  /* S[d][0] = max(0,-SL), S[d][1] = max(0,SR) at d-faces */
  double *dv[3][3]; /* dv[d][c] = velocity-component-c slope along d, per cell */
  double *cdt;      /* C_dt                                              */
  double *ex, *ey, *ez;
  double *fb[3][2]; /* face EMFs from the B fluxes (ct_emf.c:131-187):   */
                    /*   x: ezi, eyi; y: ezj, exj; z: eyk, exk           */
  double *fm[3];    /* mass flux at d-faces (UCT_CONTACT's sign)         */
  double mach;
} Work;

/* CT_EMF_AVERAGE labels of ct.h:18-21 */
enum { EMF_ARITHMETIC = 1, EMF_UCT0 = 2, EMF_UCT_CONTACT = 3, EMF_UCT_HLL = 4 };
static int g_emf = EMF_UCT_HLL;
int ora_set_emf (int e)
{
  if (e < EMF_ARITHMETIC || e > EMF_UCT_HLL) return -1;
  g_emf = e;
  return 0;
}

/* Sweep along direction d over every row of the periodic box: PLM, HLLD,
   flux difference into U, EMF face data, C_dt (stage 1). */
static void sweep_dir (const Box *g, int d, const double *V,
                       const double *bface, double dt, int stage, Work *w)
{
  const int nn = (d == 0 ? g->nx : d == 1 ? g->ny : g->nz);
  const int nt = (d == 0 ? g->ny : g->nx);
  const int nb = (d == 2 ? g->ny : g->nz);
  const double dtdx = dt/g->dx[d];
  const double inv_dl = 1.0/g->dx[d];
  const Perm p = perm_of (d);
  double mach_all = 0.0;
  long r;

  #pragma omp parallel for schedule(static) reduction(max:mach_all)
  for (r = 0; r < (long)nt*nb; r++){
    int t = (int)(r % nt), b = (int)(r / nt);
    int q, nv;
    double *row = malloc ((size_t)(nn + 4)*NV*sizeof(double));
    double *vp  = malloc ((size_t)(nn + 2)*NV*sizeof(double));
    double *vm  = malloc ((size_t)(nn + 2)*NV*sizeof(double));
    double *fl  = malloc ((size_t)(nn + 1)*NV*sizeof(double));
    double *cm  = malloc ((size_t)(nn + 1)*sizeof(double));
    double *pr  = malloc ((size_t)(nn + 1)*sizeof(double));
    double mach = 0.0;

    /* gather cells q = -2 .. nn+1 (row index q+2) */
    for (q = -2; q < nn + 2; q++){
      int i, j, k;
      long c;
      if (d == 0){ i = q; j = t; k = b; }
      else if (d == 1){ i = t; j = q; k = b; }
      else { i = t; j = b; k = q; }
      c = IDX (g, i, j, k);
      for (nv = 0; nv < NV; nv++) row[(q + 2)*NV + nv] = V[nv*g->n + c];
    }
    /* PLM on cells q = -1 .. nn (vp/vm index q+1) */
    for (q = -1; q <= nn; q++){
      plm_mc (row + (q + 1)*NV, row + (q + 2)*NV, row + (q + 3)*NV,
              vp + (q + 1)*NV, vm + (q + 1)*NV);
    }
    /* Riemann problems at interfaces q+1/2, q = -1 .. nn-1 (index q+1) */
    for (q = -1; q < nn; q++){
      double vL[NV], vR[NV], bn;
      int i, j, k;
      HLLDOut o;
      if (d == 0){ i = q; j = t; k = b; }
      else if (d == 1){ i = t; j = q; k = b; }
      else { i = t; j = b; k = q; }
      bn = bface[IDX (g, i, j, k)];
      memcpy (vL, vp + (q + 1)*NV, sizeof(vL));
      memcpy (vR, vm + (q + 2)*NV, sizeof(vR));
      vL[p.bn] = bn;
      vR[p.bn] = bn;
      riemann (vL, vR, p, g->gamma, &o);
      if (o.mach > mach) mach = MAXD(o.mach, mach);
      for (nv = 0; nv < NV; nv++) fl[(q + 1)*NV + nv] = o.flux[nv];
      pr[q + 1] = o.press;
      cm[q + 1] = o.cmax;
      if (q >= 0){
        long c = IDX (g, i, j, k);
        w->S[d][0][c] = MAXD(0.0, -o.SL);   /* ct_emf.c:145-146 */
        w->S[d][1][c] = MAXD(0.0,  o.SR);
        /* ct_emf.c:131-133, 153-155, 175-176 */
        if (d == 0){ w->fb[0][0][c] = -o.flux[BX2]; w->fb[0][1][c] =  o.flux[BX3]; }
        else if (d == 1){ w->fb[1][0][c] =  o.flux[BX1]; w->fb[1][1][c] = -o.flux[BX3]; }
        else { w->fb[2][0][c] = -o.flux[BX1]; w->fb[2][1][c] =  o.flux[BX2]; }
        w->fm[d][c] = o.flux[RHO];
      }
    }
    /* update cells q = 0 .. nn-1 */
    for (q = 0; q < nn; q++){
      int i, j, k;
      long c;
      if (d == 0){ i = q; j = t; k = b; }
      else if (d == 1){ i = t; j = q; k = b; }
      else { i = t; j = b; k = q; }
      c = IDX (g, i, j, k);
      {
        double rhs[NV];
        for (nv = 0; nv < NV; nv++) rhs[nv] = -dtdx*(fl[(q + 1)*NV + nv] - fl[q*NV + nv]);
        rhs[p.vn] -= dtdx*(pr[q + 1] - pr[q]);
        for (nv = 0; nv < NV; nv++) w->U[nv*g->n + c] += rhs[nv];
      }
      /* velocity slopes vp - vm (ct_stag_slopes.c:22-46) */
      w->dv[d][0][c] = vp[(q + 1)*NV + VX1] - vm[(q + 1)*NV + VX1];
      w->dv[d][1][c] = vp[(q + 1)*NV + VX2] - vm[(q + 1)*NV + VX2];
      w->dv[d][2][c] = vp[(q + 1)*NV + VX3] - vm[(q + 1)*NV + VX3];
      if (stage == 1){
        w->cdt[c] += 0.5*(cm[q] + cm[q + 1])*inv_dl;
      }
    }
    if (mach > mach_all) mach_all = mach;
    free (row); free (vp); free (vm); free (fl); free (cm); free (pr);
  }
  if (mach_all > w->mach) w->mach = mach_all;
}

/* MC_LIM2 (ct_stag_slopes.c:92-106) */
static double mc_lim2 (double dp, double dm)
{
  double dc, scrh;
  if (dp*dm < 0.0) return 0.0;
  dc   = 0.5*(dp + dm);
  scrh = 2.0*(fabs(dp) < fabs(dm) ? dp : dm);
  return (fabs(dc) < fabs(scrh) ? dc : scrh);
}

/* slope along axis a (0,1,2) of face array f at cell c */
static double fslope (const Box *g, const double *f, int i, int j, int k, int a)
{
  long cp, c0 = IDX (g, i, j, k), cm;
  if (a == 0){ cp = IDX (g, i+1, j, k); cm = IDX (g, i-1, j, k); }
  else if (a == 1){ cp = IDX (g, i, j+1, k); cm = IDX (g, i, j-1, k); }
  else { cp = IDX (g, i, j, k+1); cm = IDX (g, i, j, k-1); }
  return mc_lim2 (f[cp] - f[c0], f[c0] - f[cm]);
}

/* UCT_HLL edge EMFs (ct_emf_average.c:258-350), all edges of the box. */
static void emf_uct_hll (const Box *g, const double *V, const double *bx,
                         const double *by, const double *bz, Work *w)
{
  long r;
  const double *vx = V + VX1*g->n, *vy = V + VX2*g->n, *vz = V + VX3*g->n;
  double *const *S = NULL;
  (void)S;

  #pragma omp parallel for schedule(static)
  for (r = 0; r < g->n; r++){
    int i = (int)(r % g->nx), j = (int)((r / g->nx) % g->ny), k = (int)(r / ((long)g->nx*g->ny));
    long c = r;
    long cip = IDX (g, i+1, j, k), cjp = IDX (g, i, j+1, k), ckp = IDX (g, i, j, k+1);
    long cipjp = IDX (g, i+1, j+1, k), cipkp = IDX (g, i+1, j, k+1), cjpkp = IDX (g, i, j+1, k+1);
    double a_xp, a_xm, a_yp, a_ym, bS, bN, bW, bE, eSW, eSE, eNE, eNW, e;
    const double *dvx_dx = w->dv[0][0], *dvy_dx = w->dv[0][1], *dvz_dx = w->dv[0][2];
    const double *dvx_dy = w->dv[1][0], *dvy_dy = w->dv[1][1], *dvz_dy = w->dv[1][2];
    const double *dvx_dz = w->dv[2][0], *dvy_dz = w->dv[2][1], *dvz_dz = w->dv[2][2];
    const double *SxL = w->S[0][0], *SxR = w->S[0][1];
    const double *SyL = w->S[1][0], *SyR = w->S[1][1];
    const double *SzL = w->S[2][0], *SzR = w->S[2][1];

    /* Ez at (i+1/2, j+1/2, k) */
    a_xp = MAXD(SxR[c], SxR[cjp]);
    a_xm = MAXD(SxL[c], SxL[cjp]);
    a_yp = MAXD(SyR[c], SyR[cip]);
    a_ym = MAXD(SyL[c], SyL[cip]);

    bS = bx[c]   + 0.5*fslope (g, bx, i, j, k, 1);
    bW = by[c]   + 0.5*fslope (g, by, i, j, k, 0);
    bN = bx[cjp] - 0.5*fslope (g, bx, i, j+1, k, 1);
    bE = by[cip] - 0.5*fslope (g, by, i+1, j, k, 0);

    eSW = (vy[c] + 0.5*(dvy_dx[c] + dvy_dy[c]))*bS - (vx[c] + 0.5*(dvx_dx[c] + dvx_dy[c]))*bW;
    eSE = (vy[cip] - 0.5*(dvy_dx[cip] - dvy_dy[cip]))*bS - (vx[cip] - 0.5*(dvx_dx[cip] - dvx_dy[cip]))*bE;
    eNE = (vy[cipjp] - 0.5*(dvy_dx[cipjp] + dvy_dy[cipjp]))*bN - (vx[cipjp] - 0.5*(dvx_dx[cipjp] + dvx_dy[cipjp]))*bE;
    eNW = (vy[cjp] + 0.5*(dvy_dx[cjp] - dvy_dy[cjp]))*bN - (vx[cjp] + 0.5*(dvx_dx[cjp] - dvx_dy[cjp]))*bW;

    e  =   a_xp*a_yp*eSW + a_xm*a_yp*eSE + a_xm*a_ym*eNE + a_xp*a_ym*eNW;
    e /= (a_xp + a_xm)*(a_yp + a_ym);
    e -= a_yp*a_ym*(bN - bS)/(a_yp + a_ym);
    e += a_xp*a_xm*(bE - bW)/(a_xp + a_xm);
    w->ez[c] = e;

    if (g->dims < 3) continue;

    /* Ex at (i, j+1/2, k+1/2) */
    a_xp = MAXD(SyR[c], SyR[ckp]);
    a_xm = MAXD(SyL[c], SyL[ckp]);
    a_yp = MAXD(SzR[c], SzR[cjp]);
    a_ym = MAXD(SzL[c], SzL[cjp]);

    bS = by[c]   + 0.5*fslope (g, by, i, j, k, 2);
    bW = bz[c]   + 0.5*fslope (g, bz, i, j, k, 1);
    bN = by[ckp] - 0.5*fslope (g, by, i, j, k+1, 2);
    bE = bz[cjp] - 0.5*fslope (g, bz, i, j+1, k, 1);

    eSW = (vz[c] + 0.5*(dvz_dy[c] + dvz_dz[c]))*bS - (vy[c] + 0.5*(dvy_dy[c] + dvy_dz[c]))*bW;
    eSE = (vz[cjp] - 0.5*(dvz_dy[cjp] - dvz_dz[cjp]))*bS - (vy[cjp] - 0.5*(dvy_dy[cjp] - dvy_dz[cjp]))*bE;
    eNE = (vz[cjpkp] - 0.5*(dvz_dy[cjpkp] + dvz_dz[cjpkp]))*bN - (vy[cjpkp] - 0.5*(dvy_dy[cjpkp] + dvy_dz[cjpkp]))*bE;
    eNW = (vz[ckp] + 0.5*(dvz_dy[ckp] - dvz_dz[ckp]))*bN - (vy[ckp] + 0.5*(dvy_dy[ckp] - dvy_dz[ckp]))*bW;

    e  =   a_xp*a_yp*eSW + a_xm*a_yp*eSE + a_xm*a_ym*eNE + a_xp*a_ym*eNW;
    e /= (a_xp + a_xm)*(a_yp + a_ym);
    e -= a_yp*a_ym*(bN - bS)/(a_yp + a_ym);
    e += a_xp*a_xm*(bE - bW)/(a_xp + a_xm);
    w->ex[c] = e;

    /* Ey at (i+1/2, j, k+1/2); the stencil pattern follows
       ct_emf_average.c:325-348 literally (note eSE/eNW offsets). */
    a_xp = MAXD(SzR[c], SzR[cip]);
    a_xm = MAXD(SzL[c], SzL[cip]);
    a_yp = MAXD(SxR[c], SxR[ckp]);
    a_ym = MAXD(SxL[c], SxL[ckp]);

    bS = bz[c]   + 0.5*fslope (g, bz, i, j, k, 0);
    bW = bx[c]   + 0.5*fslope (g, bx, i, j, k, 2);
    bN = bz[cip] - 0.5*fslope (g, bz, i+1, j, k, 0);
    bE = bx[ckp] - 0.5*fslope (g, bx, i, j, k+1, 2);

    eSW = (vx[c] + 0.5*(dvx_dz[c] + dvx_dx[c]))*bS - (vz[c] + 0.5*(dvz_dz[c] + dvz_dx[c]))*bW;
    eSE = (vx[ckp] - 0.5*(dvx_dz[ckp] - dvx_dx[ckp]))*bS - (vz[ckp] - 0.5*(dvz_dz[ckp] - dvz_dx[ckp]))*bE;
    eNE = (vx[cipkp] - 0.5*(dvx_dz[cipkp] + dvx_dx[cipkp]))*bN - (vz[cipkp] - 0.5*(dvz_dz[cipkp] + dvz_dx[cipkp]))*bE;
    eNW = (vx[cip] + 0.5*(dvx_dz[cip] - dvx_dx[cip]))*bN - (vz[cip] + 0.5*(dvz_dz[cip] - dvz_dx[cip]))*bW;

    e  =   a_xp*a_yp*eSW + a_xm*a_yp*eSE + a_xm*a_ym*eNE + a_xp*a_ym*eNW;
    e /= (a_xp + a_xm)*(a_yp + a_ym);
    e -= a_yp*a_ym*(bN - bS)/(a_yp + a_ym);
    e += a_xp*a_xm*(bE - bW)/(a_xp + a_xm);
    w->ey[c] = e;
  }
}

/* Cell-centred inductive field -v x B of cell (i,j,k), component a
   (CT_ComputeCenterEMF, ct_emf.c:348-388). */
static double ecell (const Box *g, const double *V, int a, int i, int j, int k)
{
  long c = IDX (g, i, j, k);
  double vx = V[VX1*g->n + c], vy = V[VX2*g->n + c], vz = V[VX3*g->n + c];
  double Bx = V[BX1*g->n + c], By = V[BX2*g->n + c], Bz = V[BX3*g->n + c];
  if (a == 0) return vz*By - vy*Bz;
  if (a == 1) return vx*Bz - vz*Bx;
  return vy*Bx - vx*By;
}

/* UCT_CONTACT upwind sign of the mass flux at the d-face of (i,j,k)
   (ct_emf.c:136-141) */
static int svel (const Box *g, const Work *w, int d, int i, int j, int k)
{
  double f = w->fm[d][IDX (g, i, j, k)];
  return f > 1.e-6 ? 1 : (f < -1.e-6 ? -1 : 0);
}

/* Face EMF (stored array a of direction d) at (i,j,k) minus the cell EMF
   component e there: the P-terms of ct_emf_average.c:136-147; the M-terms
   are the cell EMF minus the face EMF one cell below along the face's own
   direction. */
#define FACE(d, a, i, j, k) (w->fb[d][a][IDX (g, (i), (j), (k))])
#define DP(d, a, e, i, j, k)  (FACE (d, a, i, j, k) - ecell (g, V, e, i, j, k))

/* Edge EMFs of the ARITHMETIC, UCT0 and UCT_CONTACT averages
   (CT_ComputeEMF, ct_emf.c:216-285; CT_EMF_ArithmeticAverage,
   ct_emf_average.c:13-44; CT_EMF_IntegrateToCorner, ct_emf_average.c:46-177).
   Face arrays: ezi = fb[0][0], eyi = fb[0][1], ezj = fb[1][0], exj = fb[1][1],
   eyk = fb[2][0], exk = fb[2][1].  UCT_CONTACT's scatter loop over cells
   (k, j, i ascending) is restated per edge: the edge (i,j,k) receives the
   two P-terms of its own cell, then the M-term of the cell after it along
   the faster loop index, then that of the next one along the slower index,
   which is the order in which the reference loop adds them. */
static void emf_face_average (const Box *g, const double *V, Work *w)
{
  long r;
  const int mode = g_emf;
  #pragma omp parallel for schedule(static)
  for (r = 0; r < g->n; r++){
    int i = (int)(r % g->nx), j = (int)((r / g->nx) % g->ny), k = (int)(r / ((long)g->nx*g->ny));
    double ez, ex, ey;

    if (mode == EMF_UCT0){
      /* ct_emf.c:261-283: face EMFs doubled minus the mean cell EMF */
#define U0(d, a, e, i2, j2, k2, di, dj, dk) \
      (FACE (d, a, i2, j2, k2)*2.0 - 0.5*(ecell (g, V, e, i2, j2, k2) \
                                         + ecell (g, V, e, (i2) + (di), (j2) + (dj), (k2) + (dk))))
      ez = 0.25*(  U0 (0, 0, 2, i, j, k, 1, 0, 0) + U0 (0, 0, 2, i, j + 1, k, 1, 0, 0)
                 + U0 (1, 0, 2, i, j, k, 0, 1, 0) + U0 (1, 0, 2, i + 1, j, k, 0, 1, 0));
      w->ez[r] = ez;
      if (g->dims < 3) continue;
      ex = 0.25*(  U0 (2, 1, 0, i, j, k, 0, 0, 1) + U0 (2, 1, 0, i, j + 1, k, 0, 0, 1)
                 + U0 (1, 1, 0, i, j, k, 0, 1, 0) + U0 (1, 1, 0, i, j, k + 1, 0, 1, 0));
      ey = 0.25*(  U0 (0, 1, 1, i, j, k, 1, 0, 0) + U0 (0, 1, 1, i, j, k + 1, 1, 0, 0)
                 + U0 (2, 0, 1, i, j, k, 0, 0, 1) + U0 (2, 0, 1, i + 1, j, k, 0, 0, 1));
#undef U0
      w->ex[r] = ex;
      w->ey[r] = ey;
      continue;
    }

    /* arithmetic sum of the four face EMFs (weight 0.25, or 1 before the
       corner integration of UCT_CONTACT) */
    {
      const double wgt = mode == EMF_ARITHMETIC ? 0.25 : 1.0;
      ez = wgt*(FACE (0, 0, i, j, k) + FACE (0, 0, i, j + 1, k)
              + FACE (1, 0, i, j, k) + FACE (1, 0, i + 1, j, k));
      if (g->dims == 3){
        ex = wgt*(FACE (2, 1, i, j, k) + FACE (2, 1, i, j + 1, k)
                + FACE (1, 1, i, j, k) + FACE (1, 1, i, j, k + 1));
        ey = wgt*(FACE (0, 1, i, j, k) + FACE (0, 1, i, j, k + 1)
                + FACE (2, 0, i, j, k) + FACE (2, 0, i + 1, j, k));
      }
    }

    if (mode == EMF_UCT_CONTACT){
      int s, u;
      /* Ez: X-span P of (i,j,k), Y-span P of (i,j,k), Y-span M of (i+1,j,k),
         X-span M of (i,j+1,k) */
      s = svel (g, w, 0, i, j, k); u = s > 0 ? i : i + 1;
      ez += s == 0 ? 0.5*(DP (1, 0, 2, i, j, k) + DP (1, 0, 2, i + 1, j, k)) : DP (1, 0, 2, u, j, k);
      s = svel (g, w, 1, i, j, k); u = s > 0 ? j : j + 1;
      ez += s == 0 ? 0.5*(DP (0, 0, 2, i, j, k) + DP (0, 0, 2, i, j + 1, k)) : DP (0, 0, 2, i, u, k);
#define DM(d, a, e, i2, j2, k2, di, dj, dk) \
      (ecell (g, V, e, i2, j2, k2) - FACE (d, a, (i2) - (di), (j2) - (dj), (k2) - (dk)))
      s = svel (g, w, 1, i + 1, j, k); u = s > 0 ? j : j + 1;
      ez -= s == 0 ? 0.5*(DM (0, 0, 2, i + 1, j, k, 1, 0, 0) + DM (0, 0, 2, i + 1, j + 1, k, 1, 0, 0))
                   : DM (0, 0, 2, i + 1, u, k, 1, 0, 0);
      s = svel (g, w, 0, i, j + 1, k); u = s > 0 ? i : i + 1;
      ez -= s == 0 ? 0.5*(DM (1, 0, 2, i, j + 1, k, 0, 1, 0) + DM (1, 0, 2, i + 1, j + 1, k, 0, 1, 0))
                   : DM (1, 0, 2, u, j + 1, k, 0, 1, 0);
      ez *= 0.25;
      if (g->dims == 3){
        /* Ex: Y-span P (dEx/dz), Z-span P (dEx/dy), Z-span M of (i,j+1,k),
           Y-span M of (i,j,k+1) */
        s = svel (g, w, 1, i, j, k); u = s > 0 ? j : j + 1;
        ex += s == 0 ? 0.5*(DP (2, 1, 0, i, j, k) + DP (2, 1, 0, i, j + 1, k)) : DP (2, 1, 0, i, u, k);
        s = svel (g, w, 2, i, j, k); u = s > 0 ? k : k + 1;
        ex += s == 0 ? 0.5*(DP (1, 1, 0, i, j, k) + DP (1, 1, 0, i, j, k + 1)) : DP (1, 1, 0, i, j, u);
        s = svel (g, w, 2, i, j + 1, k); u = s > 0 ? k : k + 1;
        ex -= s == 0 ? 0.5*(DM (1, 1, 0, i, j + 1, k, 0, 1, 0) + DM (1, 1, 0, i, j + 1, k + 1, 0, 1, 0))
                     : DM (1, 1, 0, i, j + 1, u, 0, 1, 0);
        s = svel (g, w, 1, i, j, k + 1); u = s > 0 ? j : j + 1;
        ex -= s == 0 ? 0.5*(DM (2, 1, 0, i, j, k + 1, 0, 0, 1) + DM (2, 1, 0, i, j + 1, k + 1, 0, 0, 1))
                     : DM (2, 1, 0, i, u, k + 1, 0, 0, 1);
        ex *= 0.25;
        /* Ey: X-span P (dEy/dz), Z-span P (dEy/dx), Z-span M of (i+1,j,k),
           X-span M of (i,j,k+1) */
        s = svel (g, w, 0, i, j, k); u = s > 0 ? i : i + 1;
        ey += s == 0 ? 0.5*(DP (2, 0, 1, i, j, k) + DP (2, 0, 1, i + 1, j, k)) : DP (2, 0, 1, u, j, k);
        s = svel (g, w, 2, i, j, k); u = s > 0 ? k : k + 1;
        ey += s == 0 ? 0.5*(DP (0, 1, 1, i, j, k) + DP (0, 1, 1, i, j, k + 1)) : DP (0, 1, 1, i, j, u);
        s = svel (g, w, 2, i + 1, j, k); u = s > 0 ? k : k + 1;
        ey -= s == 0 ? 0.5*(DM (0, 1, 1, i + 1, j, k, 1, 0, 0) + DM (0, 1, 1, i + 1, j, k + 1, 1, 0, 0))
                     : DM (0, 1, 1, i + 1, j, u, 1, 0, 0);
        s = svel (g, w, 0, i, j, k + 1); u = s > 0 ? i : i + 1;
        ey -= s == 0 ? 0.5*(DM (2, 0, 1, i, j, k + 1, 0, 0, 1) + DM (2, 0, 1, i + 1, j, k + 1, 0, 0, 1))
                     : DM (2, 0, 1, u, j, k + 1, 0, 0, 1);
        ey *= 0.25;
      }
#undef DM
    }
    w->ez[r] = ez;
    if (g->dims == 3){
      w->ex[r] = ex;
      w->ey[r] = ey;
    }
  }
}
#undef FACE
#undef DP

/* CT_Update (ct_update.c:79-217), Cartesian; faces updated in place. */
static void ct_update (const Box *g, double *bx, double *by, double *bz,
                       double dt, const Work *w)
{
  long r;
  const double dtx = dt/g->dx[0], dty = dt/g->dx[1], dtz = dt/g->dx[2];
  double *nbx = malloc (g->n*sizeof(double));
  double *nby = malloc (g->n*sizeof(double));
  double *nbz = malloc (g->n*sizeof(double));

  #pragma omp parallel for schedule(static)
  for (r = 0; r < g->n; r++){
    int i = (int)(r % g->nx), j = (int)((r / g->nx) % g->ny), k = (int)(r / ((long)g->nx*g->ny));
    long cim = IDX (g, i-1, j, k), cjm = IDX (g, i, j-1, k), ckm = IDX (g, i, j, k-1);
    double rx, ry, rz;
    if (g->dims == 3){
      rx = 0.0 - dty*(w->ez[r] - w->ez[cjm]) + dtz*(w->ey[r] - w->ey[ckm]);
      ry = dtx*(w->ez[r] - w->ez[cim]) - dtz*(w->ex[r] - w->ex[ckm]);
      rz = - dtx*(w->ey[r] - w->ey[cim]) + dty*(w->ex[r] - w->ex[cjm]);
      nbz[r] = bz[r] + rz;
    }else{
      rx = 0.0 - dty*(w->ez[r] - w->ez[cjm]);
      ry = dtx*(w->ez[r] - w->ez[cim]);
      nbz[r] = bz[r];
    }
    nbx[r] = bx[r] + rx;
    nby[r] = by[r] + ry;
  }
  memcpy (bx, nbx, g->n*sizeof(double));
  memcpy (by, nby, g->n*sizeof(double));
  memcpy (bz, nbz, g->n*sizeof(double));
  free (nbx); free (nby); free (nbz);
}

static void work_alloc (const Box *g, Work *w)
{
  int d, c;
  memset (w, 0, sizeof(*w));
  w->U = malloc (NV*g->n*sizeof(double));
  for (d = 0; d < 3; d++){
    for (c = 0; c < 2; c++) w->S[d][c] = calloc (g->n, sizeof(double));
    for (c = 0; c < 3; c++) w->dv[d][c] = calloc (g->n, sizeof(double));
  }
  for (d = 0; d < 3; d++){
    for (c = 0; c < 2; c++) w->fb[d][c] = calloc (g->n, sizeof(double));
    w->fm[d] = calloc (g->n, sizeof(double));
  }
  w->cdt = calloc (g->n, sizeof(double));
  w->ex = calloc (g->n, sizeof(double));
  w->ey = calloc (g->n, sizeof(double));
  w->ez = calloc (g->n, sizeof(double));
}

static void work_free (Work *w)
{
  int d, c;
  free (w->U); free (w->cdt); free (w->ex); free (w->ey); free (w->ez);
  for (d = 0; d < 3; d++){
    for (c = 0; c < 2; c++) free (w->S[d][c]);
    for (c = 0; c < 3; c++) free (w->dv[d][c]);
    for (c = 0; c < 2; c++) free (w->fb[d][c]);
    free (w->fm[d]);
  }
}

/* One UpdateStage + CT update on (V, faces) accumulating into w->U. */
static void update_stage (const Box *g, const double *V, double *bx, double *by,
                          double *bz, double dt, int stage, Work *w)
{
  const double *bf[3] = {bx, by, bz};
  int d;
  g_two_comp = (g->dims == 2);
  if (stage == 1) memset (w->cdt, 0, g->n*sizeof(double));
  for (d = 0; d < g->dims; d++) sweep_dir (g, d, V, bf[d], dt, stage, w);
  if (g_emf == EMF_UCT_HLL) emf_uct_hll (g, V, bx, by, bz, w);
  else emf_face_average (g, V, w);
  ct_update (g, bx, by, bz, dt, w);
}

/* face average into U's B slots (ct_field_average.c:64-66) */
static void average_b (const Box *g, double *U, const double *bx,
                       const double *by, const double *bz)
{
  long r;
  #pragma omp parallel for schedule(static)
  for (r = 0; r < g->n; r++){
    int i = (int)(r % g->nx), j = (int)((r / g->nx) % g->ny), k = (int)(r / ((long)g->nx*g->ny));
    U[BX1*g->n + r] = 0.5*(bx[r] + bx[IDX (g, i-1, j, k)]);
    U[BX2*g->n + r] = 0.5*(by[r] + by[IDX (g, i, j-1, k)]);
    if (g->dims == 3) U[BX3*g->n + r] = 0.5*(bz[r] + bz[IDX (g, i, j, k-1)]);
  }
}

static int cons2prim_all (const Box *g, double *U, double *V, unsigned char *flag)
{
  long r;
  const double gmm1 = g->gamma - 1.0;
  int fail = 0;
  #pragma omp parallel for schedule(static) reduction(|:fail)
  for (r = 0; r < g->n; r++){
    double u[NV], v[NV];
    int nv;
    unsigned char f = flag ? flag[r] : 0;
    for (nv = 0; nv < NV; nv++) u[nv] = U[nv*g->n + r];
    fail |= cons2prim (u, v, gmm1, &f);
    for (nv = 0; nv < NV; nv++){ U[nv*g->n + r] = u[nv]; V[nv*g->n + r] = v[nv]; }
    if (flag) flag[r] = f;
  }
  return fail;
}

static int set_box (Box *g, int nx, int ny, int nz, int dims,
                    double dx, double dy, double dz, double gamma)
{
  if (nx < 3 || ny < 3 || nz < 1 || (dims == 3 && nz < 3)) return -1;
  if (dims != 2 && dims != 3) return -1;
  if (dims == 2 && nz != 1) return -1;
  g->nx = nx; g->ny = ny; g->nz = nz; g->dims = dims;
  g->n = (long)nx*ny*nz;
  g->dx[0] = dx; g->dx[1] = dy; g->dx[2] = dz;
  g->gamma = gamma;
  return 0;
}

/* ================================================================== */
/* Public API (ctypes)                                                  */
/* ================================================================== */

int ora_num_threads (void)
{
#ifdef _OPENMP
  return omp_get_max_threads ();
#else
  return 1;
#endif
}

void ora_set_threads (int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads (n);
#else
  (void)n;
#endif
}

/* Diagnostics: when set, the conservative accumulator after each stage's
   flux update (before the face average) is copied here ([2][8][n]). */
static double *dbg_stage_u = NULL;
void ora_debug_capture (double *buf) { dbg_stage_u = buf; }

/* One RK2 step (AdvanceStep) in place.  invdt: C_dt maximum / DIMENSIONS
   from stage 1; maxmach: maximum Mach estimate over both stages.
   flag: per-cell ConsToPrim failure flags (may be NULL).
   ex/ey/ez (may be NULL): edge EMFs of stage 1, for diagnostics. */
int ora_rk2_step (int nx, int ny, int nz, int dims, double dx, double dy,
                  double dz, double gamma, double *prim, double *bx, double *by,
                  double *bz, double dt, double *invdt, double *maxmach,
                  unsigned char *flag, double *ex1, double *ey1, double *ez1)
{
  Box g;
  Work w;
  long r, n;
  int nv, fail = 0;
  double *U0, *B0[3], inv = 0.0;
  const double gmm1 = gamma - 1.0;

  if (set_box (&g, nx, ny, nz, dims, dx, dy, dz, gamma)) return -1;
  n = g.n;
  work_alloc (&g, &w);
  U0 = malloc (NV*n*sizeof(double));
  for (nv = 0; nv < 3; nv++) B0[nv] = malloc (n*sizeof(double));
  if (flag) memset (flag, 0, n);

  /* stage 1: Uc = U0 = PrimToCons(Vc), Bs0 = Vs (rk_step.c:93-99) */
  #pragma omp parallel for schedule(static)
  for (r = 0; r < n; r++){
    double v[NV], u[NV];
    int q;
    for (q = 0; q < NV; q++) v[q] = prim[q*n + r];
    prim2cons (v, u, gmm1);
    for (q = 0; q < NV; q++){ U0[q*n + r] = u[q]; w.U[q*n + r] = u[q]; }
  }
  memcpy (B0[0], bx, n*sizeof(double));
  memcpy (B0[1], by, n*sizeof(double));
  memcpy (B0[2], bz, n*sizeof(double));

  update_stage (&g, prim, bx, by, bz, dt, 1, &w);
  for (r = 0; r < n; r++) inv = MAXD(inv, w.cdt[r]);   /* update_stage.c:310-311 */
  inv /= (double)dims;
  if (dbg_stage_u) memcpy (dbg_stage_u, w.U, NV*n*sizeof(double));
  if (ex1) memcpy (ex1, w.ex, n*sizeof(double));
  if (ey1) memcpy (ey1, w.ey, n*sizeof(double));
  if (ez1) memcpy (ez1, w.ez, n*sizeof(double));
  average_b (&g, w.U, bx, by, bz);
  fail |= cons2prim_all (&g, w.U, prim, flag);

  /* stage 2 (rk_step.c:149-186) */
  update_stage (&g, prim, bx, by, bz, dt, 2, &w);
  if (dbg_stage_u) memcpy (dbg_stage_u + NV*n, w.U, NV*n*sizeof(double));
  #pragma omp parallel for schedule(static)
  for (r = 0; r < n; r++){
    int q;
    for (q = 0; q < NV; q++) w.U[q*n + r] = 0.5*U0[q*n + r] + 0.5*w.U[q*n + r];
    bx[r] = 0.5*B0[0][r] + 0.5*bx[r];
    by[r] = 0.5*B0[1][r] + 0.5*by[r];
    bz[r] = 0.5*B0[2][r] + 0.5*bz[r];
  }
  average_b (&g, w.U, bx, by, bz);
  fail |= cons2prim_all (&g, w.U, prim, flag);

  if (invdt) *invdt = inv;
  if (maxmach) *maxmach = w.mach;
  for (nv = 0; nv < 3; nv++) free (B0[nv]);
  free (U0);
  work_free (&w);
  return fail;
}

/* NextTimeStep (main.c:459-532) for the hyperbolic-only case. */
double ora_next_dt (double invdt, double cfl, double cfl_max_var, double dt_old)
{
  double dt_hyp = 1.0/invdt, dtnext;
  dt_hyp *= cfl;
  dtnext  = dt_hyp;
  dtnext  = MIND(dtnext, cfl_max_var*dt_old);
  return dtnext;
}

/* Many steps with the reference's dt control: dt_{n+1} = ora_next_dt(...).
   fixed_dt != 0 keeps dt constant.  dts (may be NULL) receives the dt used
   by each step.  Returns the number of steps whose ConsToPrim floored. */
int ora_run (int nx, int ny, int nz, int dims, double dx, double dy, double dz,
             double gamma, double *prim, double *bx, double *by, double *bz,
             int nsteps, double dt0, double cfl, double cfl_max_var,
             int fixed_dt, double *dts, double *maxmach)
{
  int s, nfail = 0;
  double dt = dt0, inv, mach;
  for (s = 0; s < nsteps; s++){
    if (dts) dts[s] = dt;
    if (ora_rk2_step (nx, ny, nz, dims, dx, dy, dz, gamma, prim, bx, by, bz, dt,
                      &inv, &mach, NULL, NULL, NULL, NULL) > 0) nfail++;
    if (maxmach) maxmach[s] = mach;
    if (!fixed_dt) dt = ora_next_dt (inv, cfl, cfl_max_var, dt);
  }
  return nfail;
}

/* ---------------- row-level entry points for pinning ------------- */

void ora_plm_row (int n, const double *v, double *vp, double *vm)
{
  int i;
  for (i = 1; i < n - 1; i++)
    plm_mc (v + (i - 1)*NV, v + i*NV, v + (i + 1)*NV, vp + i*NV, vm + i*NV);
}

void ora_hlld_rows (int dir, int n, const double *vL, const double *vR,
                    double gamma, double *flux, double *press, double *SL,
                    double *SR, double *cmax, double *mach)
{
  int i, nv;
  Perm p = perm_of (dir);
  double m = 0.0;
  g_two_comp = 0;
  for (i = 0; i < n; i++){
    HLLDOut o;
    riemann (vL + i*NV, vR + i*NV, p, gamma, &o);
    for (nv = 0; nv < NV; nv++) flux[i*NV + nv] = o.flux[nv];
    press[i] = o.press; SL[i] = o.SL; SR[i] = o.SR; cmax[i] = o.cmax;
    m = MAXD(o.mach, m);
  }
  *mach = m;
}

void ora_p2c_rows (int n, const double *v, double gamma, double *u)
{
  int i;
  for (i = 0; i < n; i++) prim2cons (v + i*NV, u + i*NV, gamma - 1.0);
}

int ora_c2p_rows (int n, const double *u, double gamma, double *v,
                  double *u_out, unsigned char *flag)
{
  int i, nv, fail = 0;
  for (i = 0; i < n; i++){
    double uu[NV];
    for (nv = 0; nv < NV; nv++) uu[nv] = u[i*NV + nv];
    flag[i] = 0;
    fail |= cons2prim (uu, v + i*NV, gamma - 1.0, flag + i);
    for (nv = 0; nv < NV; nv++) u_out[i*NV + nv] = uu[nv];
  }
  return fail;
}
