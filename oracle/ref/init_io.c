/* Problem file for driving the PLUTO reference main loop as a checker.
 *
 * PLUTO calls four user hooks from the problem file (Init, InitDomain,
 * Analysis, UserDefBoundary; Test_Problems/MHD/Orszag_Tang/init.c shows the
 * shape).  This one takes its initial state from ./state_in.bin instead of an
 * analytic formula, and writes the state from the Analysis() hook, which
 * main.c calls once per step when pluto.ini sets "analysis -1.0 1"
 * (main.c:CheckForAnalysis).
 *
 * State file layout (little-endian), shared with tests/refrun.py:
 *   char   magic[8]  = "PLTSTAT1"
 *   int32  nx, ny, nz, dims
 *   int64  step
 *   double t, dt, maxmach, spare
 *   double prim[8][nz][ny][nx]   rho, vx, vy, vz, Bx, By, Bz, p
 *   double bx[nz][ny][nx]        x-face at x_{i+1/2} of cell (i,j,k)
 *   double by[nz][ny][nx]        y-face at y_{j+1/2}
 *   double bz[nz][ny][nx]        z-face at z_{k+1/2}
 * Components absent from a 2D build (vz, Bz, bz) are written as zero and
 * must be zero on input.
 */
#include "pluto.h"

typedef struct {
  char   magic[8];
  int    nx, ny, nz, dims;
  long   step;
  double t, dt, maxmach, spare;
} StateHeader;

/* canonical slot -> PLUTO variable index (-1 = absent in this build) */
static int slot_var (int c)
{
  switch (c){
    case 0: return RHO;
    case 1: return VX1;
    case 2: return COMPONENTS >= 2 ? VX2 : -1;
    case 3: return COMPONENTS == 3 ? VX3 : -1;
    case 4: return BX1;
    case 5: return COMPONENTS >= 2 ? BX2 : -1;
    case 6: return COMPONENTS == 3 ? BX3 : -1;
    case 7: return PRS;
  }
  return -1;
}

static void die (const char *msg)
{
  fprintf (stderr, "init_io: %s\n", msg);
  exit (3);
}

#if ASSIGN_VECTOR_POTENTIAL == YES
/* Builds with ASSIGN_VECTOR_POTENTIAL YES take the vector potential from
   ./vecpot_in.bin: double ax[nz][ny][nx], ay[nz][ny][nx], az[nz][ny][nx]
   with ax at (x_i, y_{j+1/2}, z_{k+1/2}), ay at (x_{i+1/2}, y_j, z_{k+1/2}),
   az at (x_{i+1/2}, y_{j+1/2}, z_k), periodic in every index.  Startup()
   asks Init() for A at exactly those edges (startup.c:121-131) and then
   differentiates it with VectorPotentialDiff (startup.c:177-187). */
static double *g_apot = NULL;

static long edge_index (double x, int dir, int n, int *face)
{
  double t = (x - g_domBeg[dir])/((g_domEnd[dir] - g_domBeg[dir])/(double)n);
  double r = t - floor(t);
  long idx;
  *face = (r < 0.25 || r > 0.75);    /* faces sit at whole cell counts */
  idx = *face ? lround(t) - 1 : (long)floor(t);
  idx %= n;
  if (idx < 0) idx += n;
  return idx;
}

static void vecpot_at (double *us, double x1, double x2, double x3)
{
  int fx, fy, fz;
  long i, j, k, n, ncell = (long)NX1*NX2*NX3;
  if (g_apot == NULL) {
    FILE *fp = fopen ("vecpot_in.bin", "rb");
    if (fp == NULL) die ("cannot open vecpot_in.bin");
    g_apot = (double *) malloc (3*ncell*sizeof(double));
    if (fread (g_apot, sizeof(double), 3*ncell, fp) != (size_t)(3*ncell))
      die ("short vecpot_in.bin");
    fclose (fp);
  }
  i = edge_index (x1, IDIR, NX1, &fx);
  j = edge_index (x2, JDIR, NX2, &fy);
  k = edge_index (x3, KDIR, NX3, &fz);
  n = (k*NX2 + j)*NX1 + i;
  us[AX1] = us[AX2] = us[AX3] = 0.0;
  if      (!fx &&  fy &&  fz) us[AX1] = g_apot[n];
  else if ( fx && !fy &&  fz) us[AX2] = g_apot[ncell + n];
  else if ( fx &&  fy && !fz) us[AX3] = g_apot[2*ncell + n];
}
#endif

void Init (double *us, double x1, double x2, double x3)
{
  int nv;
  for (nv = 0; nv < NVAR; nv++) us[nv] = 0.0;
  us[RHO] = 1.0;
  us[PRS] = 1.0;
#if ASSIGN_VECTOR_POTENTIAL == YES
  vecpot_at (us, x1, x2, x3);
#endif
}

#ifdef PGPU_DROPIN
/* the library's AdvanceStep averages the EMF as this build's CT would */
int PGPU_SetEMFAverage (int ct_emf_average);
#endif

void InitDomain (Data *d, Grid *grid)
{
  int i, j, k, c, nv;
  long n, ncell;
  StateHeader h;
  double *buf;
  FILE *fp;
#ifdef PGPU_DROPIN
  if (PGPU_SetEMFAverage (CT_EMF_AVERAGE) != 0) die ("PGPU_SetEMFAverage failed");
#endif
  fp = fopen ("state_in.bin", "rb");
  if (fp == NULL) die ("cannot open state_in.bin");
  if (fread (&h, sizeof(h), 1, fp) != 1) die ("short header");
  if (memcmp (h.magic, "PLTSTAT1", 8) != 0) die ("bad magic");
  if (h.nx != NX1 || h.ny != NX2 || h.nz != NX3) die ("grid size mismatch");
  ncell = (long)h.nx*h.ny*h.nz;
  buf = (double *) malloc (11*ncell*sizeof(double));
  if (fread (buf, sizeof(double), 11*ncell, fp) != (size_t)(11*ncell)) die ("short data");
  fclose (fp);

  for (c = 0; c < 8; c++){
    nv = slot_var(c);
    for (k = 0; k < NX3; k++) for (j = 0; j < NX2; j++) for (i = 0; i < NX1; i++){
      n = ((long)k*NX2 + j)*NX1 + i;
      if (nv < 0) {
        if (buf[c*ncell + n] != 0.0) die ("component absent in this build is nonzero");
        continue;
      }
      d->Vc[nv][k + KBEG][j + JBEG][i + IBEG] = buf[c*ncell + n];
    }
  }

/* Faces: interior face i+1/2 sits at Vs index IBEG+i; the left boundary face
   IBEG-1 is the periodic image of the last face (it is read by the
   face-to-centre average in Startup() before Boundary() runs). */
  for (k = 0; k < NX3; k++) for (j = 0; j < NX2; j++) for (i = 0; i < NX1; i++){
    n = ((long)k*NX2 + j)*NX1 + i;
    d->Vs[BX1s][k + KBEG][j + JBEG][i + IBEG] = buf[8*ncell + n];
    if (i == NX1-1) d->Vs[BX1s][k + KBEG][j + JBEG][IBEG-1] = buf[8*ncell + n];
    #if DIMENSIONS >= 2
    d->Vs[BX2s][k + KBEG][j + JBEG][i + IBEG] = buf[9*ncell + n];
    if (j == NX2-1) d->Vs[BX2s][k + KBEG][JBEG-1][i + IBEG] = buf[9*ncell + n];
    #endif
    #if DIMENSIONS == 3
    d->Vs[BX3s][k + KBEG][j + JBEG][i + IBEG] = buf[10*ncell + n];
    if (k == NX3-1) d->Vs[BX3s][KBEG-1][j + JBEG][i + IBEG] = buf[10*ncell + n];
    #else
    if (buf[10*ncell + n] != 0.0) die ("bz must be zero in a 2D build");
    #endif
  }
  free (buf);

/* Optional ./state_low.bin: the box's own faces below the first cells
   (non-periodic boxes), replacing the images set above:
     double bx_low[nz][ny], by_low[nz][nx], bz_low[ny][nx]  (3D builds) */
  fp = fopen ("state_low.bin", "rb");
  if (fp != NULL) {
    long nlow = (long)NX3*NX2 + (long)NX3*NX1 + (long)NX2*NX1;
    double *lo = (double *) malloc (nlow*sizeof(double));
    if (DIMENSIONS != 3) die ("state_low.bin needs a 3D build");
    if (fread (lo, sizeof(double), nlow, fp) != (size_t)nlow) die ("short state_low.bin");
    fclose (fp);
    for (k = 0; k < NX3; k++) for (j = 0; j < NX2; j++)
      d->Vs[BX1s][k + KBEG][j + JBEG][IBEG-1] = lo[(long)k*NX2 + j];
    #if DIMENSIONS >= 2
    for (k = 0; k < NX3; k++) for (i = 0; i < NX1; i++)
      d->Vs[BX2s][k + KBEG][JBEG-1][i + IBEG] = lo[(long)NX3*NX2 + (long)k*NX1 + i];
    #endif
    #if DIMENSIONS == 3
    for (j = 0; j < NX2; j++) for (i = 0; i < NX1; i++)
      d->Vs[BX3s][KBEG-1][j + JBEG][i + IBEG] = lo[(long)NX3*NX2 + (long)NX3*NX1 + (long)j*NX1 + i];
    #endif
    free (lo);
  }
}

/* Dump when PLUTO_REF_DUMP lists this step ("all" dumps every call). */
static int want_dump (long step)
{
  const char *s = getenv ("PLUTO_REF_DUMP");
  char buf[4096], *tok;
  if (s == NULL) return 0;
  if (strcmp (s, "all") == 0) return 1;
  strncpy (buf, s, sizeof(buf)-1); buf[sizeof(buf)-1] = 0;
  for (tok = strtok (buf, ","); tok; tok = strtok (NULL, ",")){
    if (atol (tok) == step) return 1;
  }
  return 0;
}

void Analysis (const Data *d, Grid *grid)
{
  int i, j, k, c, nv;
  long n, ncell = (long)NX1*NX2*NX3;
  char fname[64];
  StateHeader h;
  double *buf;
  FILE *fp;

  if (!want_dump (g_stepNumber)) return;
  memset (&h, 0, sizeof(h));
  memcpy (h.magic, "PLTSTAT1", 8);
  h.nx = NX1; h.ny = NX2; h.nz = NX3; h.dims = DIMENSIONS;
  h.step = g_stepNumber; h.t = g_time; h.dt = g_dt; h.maxmach = g_maxMach;

  buf = (double *) calloc (11*ncell, sizeof(double));
  for (k = 0; k < NX3; k++) for (j = 0; j < NX2; j++) for (i = 0; i < NX1; i++){
    n = ((long)k*NX2 + j)*NX1 + i;
    for (c = 0; c < 8; c++){
      nv = slot_var(c);
      buf[c*ncell + n] = nv < 0 ? 0.0 : d->Vc[nv][k + KBEG][j + JBEG][i + IBEG];
    }
    buf[8*ncell + n] = d->Vs[BX1s][k + KBEG][j + JBEG][i + IBEG];
    #if DIMENSIONS >= 2
    buf[9*ncell + n] = d->Vs[BX2s][k + KBEG][j + JBEG][i + IBEG];
    #endif
    #if DIMENSIONS == 3
    buf[10*ncell + n] = d->Vs[BX3s][k + KBEG][j + JBEG][i + IBEG];
    #endif
  }
  sprintf (fname, "state_%06ld.bin", g_stepNumber);
  fp = fopen (fname, "wb");
  fwrite (&h, sizeof(h), 1, fp);
  fwrite (buf, sizeof(double), 11*ncell, fp);
  fclose (fp);
  free (buf);
}

/* USERDEF sides for the tests: X1_BEG is an inflow that reads the first
   interior cell, the stage and the time; every other USERDEF side copies
   the last interior layer outward with its normal velocity halved.  The
   staggered boxes set the tangential faces (and, at X1_BEG, the normal
   faces from the boundary face IBEG-1 down), as problem files do. */
void UserDefBoundary (const Data *d, RBox *box, int side, Grid *grid)
{
  int i, j, k, nv;
  double ***b;
  double *x2 = grid->x[JDIR];

  if (side == X1_BEG){
    if (box->vpos == CENTER){
      BOX_LOOP(box,k,j,i){
        d->Vc[RHO][k][j][i] = 1.2 + 0.5*g_time;
        d->Vc[VX1][k][j][i] = 0.3;
        d->Vc[VX2][k][j][i] = 0.05*g_intStage;
        d->Vc[VX3][k][j][i] = 0.02*x2[j];
        d->Vc[BX1][k][j][i] = 0.1;
        d->Vc[BX2][k][j][i] = 0.2;
        d->Vc[BX3][k][j][i] = -0.1;
        d->Vc[PRS][k][j][i] = d->Vc[PRS][k][j][IBEG];
      }
    }else if (box->vpos == X1FACE){
      BOX_LOOP(box,k,j,i) d->Vs[BX1s][k][j][i] = 0.1;
    }else if (box->vpos == X2FACE){
      BOX_LOOP(box,k,j,i) d->Vs[BX2s][k][j][i] = 0.2;
    }else if (box->vpos == X3FACE){
      BOX_LOOP(box,k,j,i) d->Vs[BX3s][k][j][i] = -0.1 + 0.01*x2[j];
    }
    return;
  }

  /* zero gradient from the layer next to the side */
  {
    int dir = (side - X1_BEG)/2, end = (side - X1_BEG)%2;
    int src[3] = {IBEG, JBEG, KBEG};
    if (end) { src[0] = IEND; src[1] = JEND; src[2] = KEND; }
    if (box->vpos == CENTER){
      BOX_LOOP(box,k,j,i){
        int ii = dir == 0 ? src[0] : i, jj = dir == 1 ? src[1] : j, kk = dir == 2 ? src[2] : k;
        for (nv = 0; nv < NVAR; nv++) d->Vc[nv][k][j][i] = d->Vc[nv][kk][jj][ii];
        d->Vc[VX1 + dir][k][j][i] = 0.5*d->Vc[VX1 + dir][kk][jj][ii];
      }
    }else if (box->vpos - X1FACE != dir){     /* tangential faces only */
      b = d->Vs[box->vpos - X1FACE];
      BOX_LOOP(box,k,j,i){
        int ii = dir == 0 ? src[0] : i, jj = dir == 1 ? src[1] : j, kk = dir == 2 ? src[2] : k;
        b[k][j][i] = b[kk][jj][ii];
      }
    }
  }
}
