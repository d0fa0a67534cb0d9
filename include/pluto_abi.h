/* pluto_abi.h -- PLUTO's own entry points for the unsplit MHD stage,
 * implemented on the MI355X by libpluto_amd.so.
 *
 * A PLUTO 4.3 build configured for PHYSICS MHD, COMPONENTS 3 (NVAR 8),
 * DIMENSIONS 3, GEOMETRY CARTESIAN, RECONSTRUCTION LINEAR (MC_LIM),
 * TIME_STEPPING RK2, DIVB_CONTROL CONSTRAINED_TRANSPORT (UCT_HLL), ideal EOS
 * and PLUTO's own boundary types links this library in place of the objects that
 * define these functions (rk_step.o for AdvanceStep, update_stage.o,
 * hlld.o, plm_states.o, mappers.o, ct_field_average.o, set_solver.o); the
 * signatures and the semantics of arguments, globals and return values are
 * the reference's.  The structs below mirror Src/structs.h byte for byte for
 * that configuration (tests/golden/abi_layout.json holds the reference's
 * sizeof/offsetof values they are checked against).
 *
 * Globals read or written by these functions (g_dir, g_intStage, g_dt,
 * g_gamma, g_maxMach) are the host program's: the library carries weak
 * definitions that a PLUTO executable's own globals.h definitions override.
 */
#ifndef PLUTO_ABI_H
#define PLUTO_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* structs (Src/structs.h)                                             */
/* ------------------------------------------------------------------ */

struct particleNode_;
struct Particle_;
struct ForcedTurb;

/* structs.h:95-146 */
typedef struct ElectroMotiveForce {
  double ***exj, ***exk, ***eyi, ***eyk, ***ezi, ***ezj;
  signed char ***svx, ***svy, ***svz;
  int ibeg, jbeg, kbeg;
  int iend, jend, kend;
  double ***SxL, ***SxR, ***SyL, ***SyR, ***SzL, ***SzR;
  double ***ex, ***ey, ***ez;
  double ***dbx_dy, ***dbx_dz;
  double ***dby_dx, ***dby_dz;
  double ***dbz_dx, ***dbz_dy;
  double ***dvx_dx, ***dvy_dx, ***dvz_dx;
  double ***dvx_dy, ***dvy_dy, ***dvz_dy;
  double ***dvx_dz, ***dvy_dz, ***dvz_dz;
} EMF;

struct timeStep_;

/* structs.h:241-275: an index box of the ghosted arrays, handed to the
   program's UserDefBoundary() */
typedef struct RBox_ {
  int ibeg, iend, jbeg, jend, kbeg, kend;
  int di, dj, dk;
  int vpos;                     /* CENTER 0, X1FACE 1, X2FACE 2, X3FACE 3 */
  int *n, *t, *b;
  int *nbeg, *nend, *tbeg, *tend, *bbeg, *bend;
} RBox;

/* structs.h:29-88 */
typedef struct Data_ {
  double ****Vc;            /* [NVAR][NX3_TOT][NX2_TOT][NX1_TOT] primitives  */
  double ****Uc;            /* [NX3_TOT][NX2_TOT][NX1_TOT][NVAR] conservative */
  double ****Vs;            /* [DIMENSIONS] staggered faces, index -1 valid  */
  double ****Vuser;
  double ***Ax1, ***Ax2, ***Ax3;
  double ****J;
  double ***Tc;
  unsigned char ***flag;
  struct particleNode_ *PHead;
  double ****Fcr, ****Jcr;
  double ***qcr;
  double ****Ecr, ****Vdust, ****Fdust;
  struct Particle_ **pstr;
  double ***Ex1, ***Ex2, ***Ex3;
  struct ElectroMotiveForce *emf;
  struct timeStep_ *Dts;
  struct ForcedTurb *Ft;
  char fill[78];
} Data;

/* structs.h:166-219 */
typedef struct Grid_ {
  double xbeg[3], xend[3];
  double xbeg_glob[3], xend_glob[3];
  double *x[3], *x_glob[3];
  double *xr[3], *xr_glob[3];
  double *xl[3], *xl_glob[3];
  double *dx[3], *dx_glob[3];
  double *xgc[3];
  double ***dV;
  double ***A[3];
  double **dx_dl[3];
  double *rt, *sp, *s, *dmu;
  double *inv_dx[3];
  double *inv_dxi[3];
  double dl_min[3];
  int np_tot_glob[3], np_int_glob[3];
  int np_tot[3], np_int[3];
  int nghost[3];
  int lbound[3], rbound[3];
  int gbeg[3], gend[3];
  int beg[3], end[3];
  int lbeg[3], lend[3];
  int uniform[3];
  int nproc[3];
  int rank_coord[3];
  int level;
  char fill[376];
} Grid;

/* structs.h:299-319 */
typedef struct State_ {
  double **v, **u, **flux, **fluxCR, **lambda, **Bbck;
  double *prs, *a2, *cw, *h;
  double **J, **cCR, **Fcr;
  double ***Lp, ***Rp;
  char fill[8];
} State;

/* structs.h:328-347 */
typedef struct Sweep_ {
  double **vn, **flux, **tc_flux;
  double *lmax;
  double **src, **rhs;
  double *press, *bn, *SL, *SR;
  unsigned char *flag;
  State stateL, stateR, stateC;
  char fill[40];
} Sweep;

/* structs.h:389-412 */
typedef struct timeStep_ {
  double *cmax;
  double invDt_hyp, invDt_par, invDt_particles, dt_cool;
  double cfl, cfl_par, rmax_par;
  double clock_particles, clock_particles_bound;
  double clock_hyp, clock_par, clock_cooling, clock_tot;
  int Nsub_particles, Nsts, Nrkc, Nrkl;
} timeStep;

typedef double ****Data_Arr;

/* pluto.h:541 */
typedef void Riemann_Solver (const Sweep *, int, int, double *, Grid *);

/* ------------------------------------------------------------------ */
/* entry points                                                        */
/* ------------------------------------------------------------------ */

/* prototypes.h:5, Time_Stepping/rk_step.c:27.  One RK2 step of the whole
   box on the GPU.  Periodic on all sides: Vc, Vs (interior) in, Vc, Vs
   (interior, plus the left boundary faces) and flag out.  With OUTFLOW,
   REFLECTIVE, AXISYMMETRIC or EQTSYMMETRIC sides (any mix with PERIODIC):
   the whole ghosted Vc, Vs in, Boundary()'s ghost fill on the device before
   each stage (rk_step.c:81,151), Vc and flag out over the domain and Vs over
   the CT_Update ranges.  Dts->invDt_hyp and g_maxMach set as by
   UpdateStage/HLL_Speed.  Uses g_dt and the EMF average set by
   PGPU_SetEMFAverage.  Returns 0, or nonzero (and changes
   nothing) when the configuration is not the supported one (Riemann not
   HLLD_Solver, HLL_Solver, HLLC_Solver, LF_Solver or Roe_Solver,
   SHEARING sides, USERDEF sides without a UserDefBoundary, a domain split
   over processes, non-uniform grid or face areas) or the device fails.
   On a periodic box the arrays travel through pinned host buffers: each
   array is packed and its upload enqueued before the next is packed, and
   each download is unpacked as soon as it lands. */
int  AdvanceStep (Data *d, Riemann_Solver *Riemann, timeStep *Dts, Grid *grid);

/* Where the time of the periodic AdvanceStep calls goes (a measurement hook
   with no PLUTO counterpart), in seconds summed over the calls since the
   last reset: pack and unpack are host time spent copying between PLUTO's
   arrays and the pinned buffers; h2d, step and d2h are device-event
   intervals (uploads, the RK2 step, downloads); total is the wall time of
   the calls.  pack overlaps h2d and unpack overlaps d2h, so the parts add
   up to more than total. */
typedef struct pgpu_dropin_times {
  double pack, h2d, step, d2h, unpack, total;
  long long calls;
} pgpu_dropin_times;
int  PGPU_DropinTimes (pgpu_dropin_times *out, int reset);

/* prototypes.h:190-191, Time_Stepping/update_stage.c:37.  One stage:
   UU += dt*R(Vc) over the domain in x, y, z order, Vs updated by CT with the
   UCT_HLL EMF; with g_intStage == 1 also Dts->invDt_hyp.  The magnetic
   components of UU are not touched (the reference adds flux differences to
   them that CT_AverageMagneticField then overwrites); aflux must be NULL
   (AMR refluxing is not supported); d->emf is not filled.  With
   non-periodic sides the stage reads the ghost zones the caller's
   Boundary() filled and writes UU over the domain and Vs over the
   CT_Update ranges. */
void UpdateStage (Data *d, Data_Arr UU, double **aflux, Riemann_Solver *Riemann,
                  double dt, timeStep *Dts, Grid *grid);

/* MHD/hlld.c:44.  HLLD at interfaces beg..end of a sweep along g_dir from
   sweep->stateL.v and stateR.v; writes sweep->flux (all NVAR components),
   sweep->press, sweep->SL, sweep->SR, cmax[] and raises g_maxMach.  The
   conservative states are recomputed from the primitive ones. */
void HLLD_Solver (const Sweep *sweep, int beg, int end, double *cmax, Grid *grid);
/* MHD/hll.c:30.  The HLL solver on the same row layout. */
void HLL_Solver (const Sweep *sweep, int beg, int end, double *cmax, Grid *grid);
/* MHD/hllc.c:32.  The HLLC solver on the same row layout. */
void HLLC_Solver (const Sweep *sweep, int beg, int end, double *cmax, Grid *grid);
/* MHD/tvdlf.c:37.  The Lax-Friedrichs (Rusanov) solver, PLUTO's "tvdlf". */
void LF_Solver (const Sweep *sweep, int beg, int end, double *cmax, Grid *grid);
/* MHD/roe.c:53.  The Roe solver (Cargo & Gallice), PLUTO's "roe"; rows of
   the three-component build. */
void Roe_Solver (const Sweep *sweep, int beg, int end, double *cmax, Grid *grid);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, serving loops, harnesses that forbid
   hipMalloc): the Riemann solver rows of HLLD_Solver, HLL_Solver,
   HLLC_Solver, LF_Solver or Roe_Solver (solver = PGPU_SOLVER_HLLD, _HLL,
   _HLLC, _TVDLF, _ROE of pluto_amd.h), with the same inputs, outputs and
   results, staged through a fixed device array of PGPU_STATIC_ROWS_MAX rows
   (1.9 MB per GPU, kept until the process ends) and fixed host arrays, on
   the null stream.  Calls no allocating routine.  Returns 0; -1 with
   nothing computed for more than PGPU_STATIC_ROWS_MAX interfaces or an
   unknown solver; -2 on a device error.  Calls share the scratch: calls
   that overlap (several host threads) race. */
#define PGPU_STATIC_ROWS_MAX 8192
int  pgpu_riemann_rows_static (int solver, const Sweep *sweep, int beg, int end,
                               double *cmax);

/* prototypes.h:186, States/plm_states.c:80.  PLM (MC limiter) of
   sweep->stateC.v: stateL.v[i] (i+1/2) and stateR.v[i-1] (i-1/2) for
   i = beg..end with the normal field taken from sweep->bn, and the
   conservative stateL.u / stateR.u. */
void States (const Sweep *sweep, int beg, int end, Grid *grid);

/* MHD/mappers.c:25 and :88. */
void PrimToCons (double **uprim, double **ucons, int ibeg, int iend);
int  ConsToPrim (double **ucons, double **uprim, int ibeg, int iend,
                 unsigned char *flag);

/* MHD/CT/ct_field_average.c:13.  UU[..][BX] = face average over the domain
   plus one layer of ghost zones. */
void CT_AverageMagneticField (double ****bf, double ****UU, Grid *grid);

/* The CT API (MHD/CT/ct.h:94-112) for the UCT_HLL average.  Row level, on
   the host: CT_StoreUpwindEMF (ct_emf.c:104-207) copies a sweep row's face
   EMFs (x: -F(By), F(Bz); y: F(Bx), -F(Bz); z: -F(Bx), F(By)), face speeds
   max(0,-SL), max(0,SR) and, through CT_StoreVelSlopes
   (ct_stag_slopes.c:5-49), the PLM velocity slopes into the EMF arrays at
   the row g_i/g_j/g_k of g_dir, and records the EMF range of g_dir.  Box
   level, on the device: CT_ComputeEMF (ct_emf.c:210-291) computes ex, ey, ez
   over the recorded ranges (the MC_LIM2 staggered slopes are computed on
   the fly, not stored in emf->db*_d*); CT_Update (ct_update.c:29-220)
   writes d->Vs = Bs + the curl terms over the reference's ranges.
   CT_Allocate (ct_emf.c:24-98) allocates the arrays these use, of the size
   the host program's NX1_TOT, NX2_TOT, NX3_TOT give. */
void CT_Allocate (EMF *emf);
void CT_StoreUpwindEMF (const Sweep *sweep, EMF *emf, int beg, int end, Grid *grid);
void CT_StoreVelSlopes (EMF *emf, const Sweep *sweep, int beg, int end);
void CT_ComputeEMF (const Data *d, Grid *grid);
void CT_Update (const Data *d, Data_Arr Bs, double dt, Grid *grid);

/* prototypes.h, boundary.c:40-315.  Fills the ghost zones of d->Vc and
   d->Vs on the sides of idim (IDIR, JDIR, KDIR or ALL_DIR = -1) on the
   device, side by side in the reference's order (X1_BEG, X1_END, X2_BEG,
   ...), for the types in grid->lbound/rbound: OUTFLOW (1), REFLECTIVE (2),
   AXISYMMETRIC (3), EQTSYMMETRIC (4), PERIODIC (5), 0 (skipped).  Outflow
   and the symmetric types also set the normal staggered field of the ghost
   faces from div.B = 0 (ct_fill_mag_field.c:38-140, with grid->A), and
   outflow the cell-centred normal field from it
   (ct_field_average.c:134-270).  USERDEF (7) sides call the program's
   UserDefBoundary (boundary.c:258-292) on the host, on the centre box and
   the three face boxes of the side with PLUTO's extents, then set the
   transverse cell-centred field (CT_AverageTransverseMagField,
   ct_field_average.c:275-457) on the host and FillMagneticField /
   CT_AverageNormalMagField on the device.  Reports an error and changes
   nothing for SHEARING sides, USERDEF sides in a program without
   UserDefBoundary, a domain split over processes, or fewer interior cells
   than ghost zones along an axis. */
void Boundary (const Data *d, int idim, Grid *grid);

/* The problem file's hook (prototypes.h; boundary.c:286-289), called for
   USERDEF sides by Boundary() and AdvanceStep().  Not defined by the
   library: a weak reference to the program's own definition. */
void UserDefBoundary (const Data *d, RBox *box, int side, Grid *grid);

/* MHD/set_solver.c:13.  Returns HLLD_Solver for "hlld", HLL_Solver for
   "hll", HLLC_Solver for "hllc", LF_Solver for "tvdlf", Roe_Solver for
   "roe", NULL otherwise. */
Riemann_Solver *SetSolver (const char *solver);

/* No counterpart in the reference: there CT_EMF_AVERAGE is a compile-time
   macro (ct.h:43-45) of the PLUTO build.  A build that links this library
   calls PGPU_SetEMFAverage(CT_EMF_AVERAGE) once (its init, say) so that
   AdvanceStep averages the EMF as its own CT would: ARITHMETIC (1), UCT0
   (2), UCT_CONTACT (3) or UCT_HLL (4, the default).  UpdateStage supports
   UCT_HLL only and reports an error otherwise.  Returns 0, or 1 for an
   unknown label. */
int PGPU_SetEMFAverage (int ct_emf_average);

/* No counterpart in the reference (its scratch lives until exit).  Frees the
   device memory the PLUTO entry points hold: the domain cached for the grid
   of AdvanceStep/UpdateStage and the scratch of the row-level functions.
   They allocate again on their next call. */
void pgpu_finalize (void);

#ifdef __cplusplus
}
#endif
#endif
