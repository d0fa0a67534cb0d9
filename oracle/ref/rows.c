/* ctypes entry points into the PLUTO reference row kernels (3D MHD build).
 *
 * Row arrays use the canonical primitive order rho, vx, vy, vz, Bx, By, Bz, p
 * ([n][8], row-major), which is also PLUTO's NVAR order for this build.
 *
 *   ref_plm_row   -> States()        States/plm_states.c:80-312
 *   ref_hlld_rows -> HLLD_Solver()   MHD/hlld.c:44-452
 *   ref_hll_rows  -> HLL_Solver()    MHD/hll.c:30-169
 *   ref_hllc_rows -> HLLC_Solver()   MHD/hllc.c:32-240
 *   ref_tvdlf_rows -> LF_Solver()    MHD/tvdlf.c:37-163
 *   ref_roe_rows  -> Roe_Solver()    MHD/roe.c:53-700
 *   ref_p2c_rows  -> PrimToCons()    MHD/mappers.c:25-86
 *   ref_c2p_rows  -> ConsToPrim()    MHD/mappers.c:88-254
 */
#include "pluto.h"
#include "globals.h"   /* defines the globals that main.c defines in a full build */

#define ROWS_MAX 4096

static Sweep g_sweep;
static Grid  g_grid[3];
static int   g_ready = 0;

static void setup (void)
{
  int d;
  if (g_ready) return;
  NMAX_POINT = ROWS_MAX + 8;
  MakeState (&g_sweep);
  memset (g_grid, 0, sizeof(g_grid));
  /* coordinates for the reference's Where(), which ConsToPrim calls when it
     applies a floor (tools.c:495-560) */
  for (d = 0; d < 3; d++){
    g_grid->x[d]  = (double *) calloc (NMAX_POINT, sizeof(double));
    g_grid->xl[d] = g_grid->x[d];
    g_grid->xr[d] = g_grid->x[d];
    g_grid->dx[d] = g_grid->x[d];
  }
  Where (-1, g_grid);
  g_ready = 1;
}

int ref_rows_max (void) { return ROWS_MAX; }

/* PLM states for cells 1..n-2 of a row of n cells; bn[i] is the staggered
   normal field at interface i+1/2.  Outputs (cells 1..n-2):
   vp[i] = state at i+1/2 (left of interface), vm[i] = state at i-1/2,
   up/um the matching conservative states. */
int ref_plm_row (int dir, int n, const double *v, const double *bn,
                 double *vp, double *vm, double *up, double *um)
{
  int i, nv;
  State *sC = &g_sweep.stateC, *sL = &g_sweep.stateL, *sR = &g_sweep.stateR;
  setup ();
  if (n < 3 || n > ROWS_MAX) return -1;
  g_dir = dir;
  SetVectorIndices (dir);
  for (i = 0; i < n; i++){
    for (nv = 0; nv < NVAR; nv++) sC->v[i][nv] = v[i*NVAR + nv];
    g_sweep.bn[i] = bn[i];
  }
  States (&g_sweep, 1, n-2, g_grid);
  for (i = 1; i <= n-2; i++){
    for (nv = 0; nv < NVAR; nv++){
      vp[i*NVAR + nv] = sL->v[i][nv];
      vm[i*NVAR + nv] = sR->v[i-1][nv];
      up[i*NVAR + nv] = sL->u[i][nv];
      um[i*NVAR + nv] = sR->u[i-1][nv];
    }
  }
  return 0;
}

/* HLLD at n interfaces given left/right primitive states. */
int ref_hlld_rows (int dir, int n, const double *vL, const double *vR,
                   double *flux, double *press, double *SL, double *SR,
                   double *cmax, double *maxmach)
{
  int i, nv;
  State *sL = &g_sweep.stateL, *sR = &g_sweep.stateR;
  setup ();
  if (n < 1 || n > ROWS_MAX) return -1;
  g_dir = dir;
  SetVectorIndices (dir);
  for (i = 0; i < n; i++){
    for (nv = 0; nv < NVAR; nv++){
      sL->v[i][nv] = vL[i*NVAR + nv];
      sR->v[i][nv] = vR[i*NVAR + nv];
    }
  }
  PrimToCons (sL->v, sL->u, 0, n-1);
  PrimToCons (sR->v, sR->u, 0, n-1);
  g_maxMach = 0.0;
  HLLD_Solver (&g_sweep, 0, n-1, cmax, g_grid);
  for (i = 0; i < n; i++){
    for (nv = 0; nv < NVAR; nv++) flux[i*NVAR + nv] = g_sweep.flux[i][nv];
    press[i] = g_sweep.press[i];
    SL[i] = g_sweep.SL[i];
    SR[i] = g_sweep.SR[i];
  }
  *maxmach = g_maxMach;
  return 0;
}

/* One of the reference's Riemann solvers on the rows of ref_hlld_rows. */
static int solver_rows (Riemann_Solver *solver, int dir, int n, const double *vL,
                        const double *vR, double *flux, double *press, double *SL,
                        double *SR, double *cmax, double *maxmach)
{
  int i, nv;
  State *sL = &g_sweep.stateL, *sR = &g_sweep.stateR;
  setup ();
  if (n < 1 || n > ROWS_MAX) return -1;
  g_dir = dir;
  SetVectorIndices (dir);
  for (i = 0; i < n; i++){
    for (nv = 0; nv < NVAR; nv++){
      sL->v[i][nv] = vL[i*NVAR + nv];
      sR->v[i][nv] = vR[i*NVAR + nv];
    }
  }
  PrimToCons (sL->v, sL->u, 0, n-1);
  PrimToCons (sR->v, sR->u, 0, n-1);
  g_maxMach = 0.0;
  solver (&g_sweep, 0, n-1, cmax, g_grid);
  for (i = 0; i < n; i++){
    for (nv = 0; nv < NVAR; nv++) flux[i*NVAR + nv] = g_sweep.flux[i][nv];
    press[i] = g_sweep.press[i];
    SL[i] = g_sweep.SL[i];
    SR[i] = g_sweep.SR[i];
  }
  *maxmach = g_maxMach;
  return 0;
}

/* HLL_Solver on the same rows (MHD/hll.c:30-169). */
int ref_hll_rows (int dir, int n, const double *vL, const double *vR,
                  double *flux, double *press, double *SL, double *SR,
                  double *cmax, double *maxmach)
{
  return solver_rows (HLL_Solver, dir, n, vL, vR, flux, press, SL, SR, cmax, maxmach);
}

/* HLLC_Solver on the same rows (MHD/hllc.c:32-240). */
int ref_hllc_rows (int dir, int n, const double *vL, const double *vR,
                   double *flux, double *press, double *SL, double *SR,
                   double *cmax, double *maxmach)
{
  return solver_rows (HLLC_Solver, dir, n, vL, vR, flux, press, SL, SR, cmax, maxmach);
}

/* LF_Solver ("tvdlf") on the same rows (MHD/tvdlf.c:37-163). */
int ref_tvdlf_rows (int dir, int n, const double *vL, const double *vR,
                    double *flux, double *press, double *SL, double *SR,
                    double *cmax, double *maxmach)
{
  return solver_rows (LF_Solver, dir, n, vL, vR, flux, press, SL, SR, cmax, maxmach);
}

/* Roe_Solver ("roe") on the same rows (MHD/roe.c:53-700). */
int ref_roe_rows (int dir, int n, const double *vL, const double *vR,
                  double *flux, double *press, double *SL, double *SR,
                  double *cmax, double *maxmach)
{
  return solver_rows (Roe_Solver, dir, n, vL, vR, flux, press, SL, SR, cmax, maxmach);
}

int ref_p2c_rows (int n, const double *v, double *u)
{
  int i, nv;
  double **vv, **uu;
  setup ();
  if (n < 1 || n > ROWS_MAX) return -1;
  vv = g_sweep.stateC.v; uu = g_sweep.stateC.u;
  for (i = 0; i < n; i++) for (nv = 0; nv < NVAR; nv++) vv[i][nv] = v[i*NVAR + nv];
  PrimToCons (vv, uu, 0, n-1);
  for (i = 0; i < n; i++) for (nv = 0; nv < NVAR; nv++) u[i*NVAR + nv] = uu[i][nv];
  return 0;
}

/* ConsToPrim; u is copied, floors applied to the copy are returned in
   u_out, failure flags in flag[]. */
int ref_c2p_rows (int n, const double *u, double *v, double *u_out,
                  unsigned char *flag)
{
  int i, nv;
  double **vv, **uu;
  setup ();
  if (n < 1 || n > ROWS_MAX) return -1;
  vv = g_sweep.stateC.v; uu = g_sweep.stateC.u;
  for (i = 0; i < n; i++){
    for (nv = 0; nv < NVAR; nv++) uu[i][nv] = u[i*NVAR + nv];
    flag[i] = 0;
  }
  g_dir = IDIR; g_i = g_j = g_k = 0;
  ConsToPrim (uu, vv, 0, n-1, flag);
  for (i = 0; i < n; i++) for (nv = 0; nv < NVAR; nv++){
    v[i*NVAR + nv] = vv[i][nv];
    u_out[i*NVAR + nv] = uu[i][nv];
  }
  return 0;
}

/* Problem-file hooks the reference objects reference; unused here. */
void Init (double *us, double x1, double x2, double x3) { }
void InitDomain (Data *d, Grid *grid) { }
void Analysis (const Data *d, Grid *grid) { }
void UserDefBoundary (const Data *d, RBox *box, int side, Grid *grid) { }
