/* AddressSanitizer / UndefinedBehaviorSanitizer run of the oracle (host
 * code only): tests/test_sanitize_cpu.py compiles this driver with
 * oracle/mhd_oracle.c under -fsanitize=address,undefined and runs it.  It
 * drives every entry point of the oracle on boxes of odd and tiny sizes
 * (2D and 3D, every Riemann solver and EMF average, floors), so an
 * out-of-bounds access or undefined arithmetic in the row and box loops
 * aborts the run. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int ora_rk2_step (int nx, int ny, int nz, int dims, double dx, double dy, double dz,
                  double gamma, double *prim, double *bx, double *by, double *bz, double dt,
                  double *invdt, double *maxmach, unsigned char *flag, double *ex1, double *ey1,
                  double *ez1);
int ora_run (int nx, int ny, int nz, int dims, double dx, double dy, double dz, double gamma,
             double *prim, double *bx, double *by, double *bz, int nsteps, double dt0,
             double cfl, double cfl_max_var, int fixed_dt, double *dts, double *maxmach);
void ora_set_solver (int s);
int ora_set_emf (int e);
void ora_set_threads (int n);
void ora_plm_row (int n, const double *v, double *vp, double *vm);
void ora_hlld_rows (int dir, int n, const double *vL, const double *vR, double gamma,
                    double *flux, double *press, double *SL, double *SR, double *cmax,
                    double *mach);
void ora_p2c_rows (int n, const double *v, double gamma, double *u);
int ora_c2p_rows (int n, const double *u, double gamma, double *v, double *u_out,
                  unsigned char *flag);

static unsigned long long rng = 88172645463325252ull;
static double urand (void)
{
  rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
  return (double)(rng >> 11) * (1.0 / 9007199254740992.0);
}

/* a state with positive density and pressure and a face field that is the
   curl of a random edge potential (periodic), so div.B = 0 */
static void make_state (int nx, int ny, int nz, int dims, double *prim, double *bx,
                        double *by, double *bz)
{
  long n = (long)nx*ny*nz, c;
  double *ax = malloc (n*sizeof(double)), *ay = malloc (n*sizeof(double)),
         *az = malloc (n*sizeof(double));
  int i, j, k;
  for (c = 0; c < n; c++){
    prim[0*n + c] = 1.0 + 0.3*urand ();
    prim[1*n + c] = urand () - 0.5;
    prim[2*n + c] = urand () - 0.5;
    prim[3*n + c] = dims == 3 ? urand () - 0.5 : 0.0;
    prim[7*n + c] = 1.0 + 0.3*urand ();
    ax[c] = dims == 3 ? 0.1*urand () : 0.0;
    ay[c] = dims == 3 ? 0.1*urand () : 0.0;
    az[c] = 0.1*urand ();
  }
  for (k = 0; k < nz; k++) for (j = 0; j < ny; j++) for (i = 0; i < nx; i++){
    long c0 = ((long)k*ny + j)*nx + i;
    long cjm = ((long)k*ny + (j + ny - 1)%ny)*nx + i;
    long cim = ((long)k*ny + j)*nx + (i + nx - 1)%nx;
    long ckm = ((long)((k + nz - 1)%nz)*ny + j)*nx + i;
    bx[c0] = (az[c0] - az[cjm])*ny - (dims == 3 ? (ay[c0] - ay[ckm])*nz : 0.0);
    by[c0] = -(az[c0] - az[cim])*nx + (dims == 3 ? (ax[c0] - ax[ckm])*nz : 0.0);
    bz[c0] = dims == 3 ? (ay[c0] - ay[cim])*nx - (ax[c0] - ax[cjm])*ny : 0.0;
  }
  for (k = 0; k < nz; k++) for (j = 0; j < ny; j++) for (i = 0; i < nx; i++){
    long c0 = ((long)k*ny + j)*nx + i;
    prim[4*n + c0] = 0.5*(bx[c0] + bx[((long)k*ny + j)*nx + (i + nx - 1)%nx]);
    prim[5*n + c0] = 0.5*(by[c0] + by[((long)k*ny + (j + ny - 1)%ny)*nx + i]);
    prim[6*n + c0] = dims == 3 ? 0.5*(bz[c0] + bz[((long)((k + nz - 1)%nz)*ny + j)*nx + i]) : 0.0;
  }
  free (ax); free (ay); free (az);
}

static int box (int nx, int ny, int nz, int dims, int solver, int emf, int steps)
{
  long n = (long)nx*ny*nz;
  double *prim = malloc (8*n*sizeof(double)), *bx = malloc (n*sizeof(double)),
         *by = malloc (n*sizeof(double)), *bz = malloc (n*sizeof(double));
  double *ex = malloc (n*sizeof(double)), *ey = malloc (n*sizeof(double)),
         *ez = malloc (n*sizeof(double));
  double inv, mach, dts[8], machs[8];
  unsigned char *flag = malloc (n);
  long c;
  int bad = 0;
  make_state (nx, ny, nz, dims, prim, bx, by, bz);
  ora_set_solver (solver);
  if (ora_set_emf (emf) != 0) return 1;
  if (ora_rk2_step (nx, ny, nz, dims, 1.0/nx, 1.0/ny, 1.0/nz, 5.0/3.0, prim, bx, by, bz, 1e-3,
                    &inv, &mach, flag, ex, ey, ez) < 0) bad = 1;
  ora_run (nx, ny, nz, dims, 1.0/nx, 1.0/ny, 1.0/nz, 5.0/3.0, prim, bx, by, bz, steps, 1e-3,
           0.4, 1.1, 0, dts, machs);
  for (c = 0; c < 8*n; c++) if (!isfinite (prim[c])) bad = 1;
  free (prim); free (bx); free (by); free (bz); free (ex); free (ey); free (ez); free (flag);
  return bad;
}

int main (void)
{
  static const int shapes[][4] = {{3, 3, 3, 3}, {5, 7, 4, 3}, {9, 4, 6, 3}, {8, 8, 1, 2},
                                  {5, 3, 1, 2}, {12, 10, 9, 3}};
  int s, solver, emf, bad = 0, n = 37, i;
  double v[37*8], vp[37*8], vm[37*8], u[37*8], uo[37*8], f[37*8], pr[37], sl[37], sr[37],
         cm[37], mach;
  unsigned char fl[37];
  ora_set_threads (1);
  for (s = 0; s < (int)(sizeof(shapes)/sizeof(shapes[0])); s++)
    for (solver = 0; solver <= 4; solver++)
      for (emf = 1; emf <= 4; emf++)
        bad |= box (shapes[s][0], shapes[s][1], shapes[s][2], shapes[s][3], solver, emf,
                    solver == 0 && emf == 4 ? 4 : 1);
  /* rows, with a floor case in ConsToPrim */
  for (i = 0; i < n*8; i++) v[i] = (i % 8 == 0 || i % 8 == 7) ? 1.0 + urand () : urand () - 0.5;
  ora_plm_row (n, v, vp, vm);
  for (s = 0; s < 3; s++){
    ora_set_solver (s % 5);
    ora_hlld_rows (s, n - 1, v, v + 8, 5.0/3.0, f, pr, sl, sr, cm, &mach);
  }
  ora_p2c_rows (n, v, 5.0/3.0, u);
  u[8*3] = -1.0;
  u[8*5 + 7] = -1.0;
  ora_c2p_rows (n, u, 5.0/3.0, v, uo, fl);
  printf ("sanitized oracle run: %s\n", bad ? "non-finite state" : "ok");
  return bad;
}
