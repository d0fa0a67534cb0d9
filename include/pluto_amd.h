/* pluto_amd.h -- native C-ABI of the MI355X unsplit MHD stage.
 *
 * A "domain" is one periodic Cartesian box resident in device memory:
 * cell-centred primitives rho, vx, vy, vz, Bx, By, Bz, p (PLUTO's NVAR order
 * for 3D MHD, MHD/mod_defs.h:20-41) and the staggered face fields bx (at
 * x_{i+1/2}), by, bz, with no ghost zones (periodic neighbours are index
 * wraps).  Arrays are [var][nz][ny][nx], x fastest.
 *
 * pgpu_run() advances the domain by whole RK2 steps of PLUTO's unsplit
 * Godunov MHD stage -- PLM with the MC limiter, HLLD, constrained transport
 * with the UCT_HLL EMF, RK2 (Heun) -- entirely on the device, including the
 * CFL time-step control of NextTimeStep (main.c:389-575), so consecutive
 * steps never wait on the host.
 *
 * PLUTO's own entry points (AdvanceStep, UpdateStage, HLLD_Solver, States,
 * the CT functions) are declared in pluto_abi.h and are built on this API.
 *
 * All functions return 0 on success and a negative PGPU_E* code on failure;
 * pgpu_last_error() describes the last failure of the calling thread.
 * Domains are not thread-safe; calls on one domain must be serialised.
 */
#ifndef PLUTO_AMD_H
#define PLUTO_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGPU_OK        0
#define PGPU_EINVAL   -1   /* bad argument or unsupported configuration */
#define PGPU_EDEVICE  -2   /* HIP runtime error or no device            */
#define PGPU_ENOMEM   -3   /* device allocation failed                  */
#define PGPU_ENAN     -4   /* a stage input held a NaN: where PLUTO's
                              CheckNaN (tools.c:20-38) stops the run     */
#define PGPU_EDTSMALL -5   /* the CFL dt fell below 1e-9 of the first dt:
                              where PLUTO's NextTimeStep (main.c:534-543)
                              stops the run; later steps took dt = 0     */

/* precision of the PLM + Riemann-solver arithmetic; the conservative
   accumulation, CT update and ConsToPrim are FP64 in both */
#define PGPU_PREC_FP64      0
#define PGPU_PREC_FP32_FLUX 1
#define PGPU_PREC_FP64_FAST 2   /* FP64 flux kernels with FMA contraction and
                                   Newton division/sqrt: not bit-exact, within
                                   1e-12 relative per step (periodic HLLD +
                                   UCT_HLL, fused pipeline) */

/* kernel pipelines of a step (identical results) */
#define PGPU_PIPELINE_FUSED   0   /* flux_xy, flux_z, emf, ct+c2p per stage (default) */
#define PGPU_PIPELINE_MODULAR 1   /* sweep x/y/z, update, emf, ct, c2p per stage       */

/* Riemann solvers (PLUTO's Solver in pluto.ini, MHD/set_solver.c) */
#define PGPU_SOLVER_HLLD 0   /* HLLD_Solver, MHD/hlld.c:44-452 (default) */
#define PGPU_SOLVER_HLL  1   /* HLL_Solver,  MHD/hll.c:30-169            */
#define PGPU_SOLVER_HLLC 2   /* HLLC_Solver, MHD/hllc.c:32-240           */
#define PGPU_SOLVER_TVDLF 3  /* LF_Solver ("tvdlf"), MHD/tvdlf.c:37-163  */
#define PGPU_SOLVER_ROE   4  /* Roe_Solver ("roe"), MHD/roe.c:53-700      */

typedef struct pgpu_domain pgpu_domain;

const char *pgpu_last_error (void);
int  pgpu_device_count (int *count);

/* dims = 3 (nz >= 3) or 2 (x-y problem, nz = 1: no z sweep, only Ez,
   C_dt divided by 2 as for DIMENSIONS 2).  dx, dy, dz are the uniform cell
   widths ((xR-xL)/N, set_grid.c:399); gamma the adiabatic index. */
int  pgpu_domain_create (int nx, int ny, int nz, int dims,
                         double dx, double dy, double dz, double gamma,
                         pgpu_domain **out);
int  pgpu_domain_destroy (pgpu_domain *dom);

/* Boundary types of a bounded box (pluto.h:146-150). */
#define PGPU_BOUND_OUTFLOW       1
#define PGPU_BOUND_REFLECTIVE    2
#define PGPU_BOUND_AXISYMMETRIC  3
#define PGPU_BOUND_EQTSYMMETRIC  4
#define PGPU_BOUND_PERIODIC      5
#define PGPU_BOUND_USERDEF       7   /* host code: pgpu_set_userdef_boundary */

/* A resident 3D box with PLUTO's boundary conditions on its six sides
   (bound[]: X1_BEG, X1_END, X2_BEG, X2_END, X3_BEG, X3_END; PGPU_BOUND_*),
   held with nghost (>= 2) ghost zones per side: every stage starts with
   Boundary()'s ghost fill on the device (boundary.c:120-292, with the
   div.B = 0 ghost faces of ct_fill_mag_field.c), and the CFL, Mach,
   failure and NaN reductions keep to PLUTO's ranges.  The state arrays of
   pgpu_set_state/get_state/get_flags are the interior cells (nx*ny*nz);
   pgpu_set_state starts the face below the first cell as the image of the
   last face, pgpu_set_state_bounded takes those faces from the caller.
   An axis is periodic on both sides or on neither.
   pgpu_write_dbl/pgpu_read_dbl write and restore the interior with the
   box's own faces below the first cells (PLUTO's file layout).  Fused FP64
   pipeline only: FP32 flux, the modular pipeline, the adjoint, the
   device-pointer state calls and pgpu_state_error refuse a bounded box.  No counterpart in the reference: PLUTO keeps the box on the
   host and calls Boundary() between stages (rk_step.c:81,151). */
int  pgpu_domain_create_bounded (int nx, int ny, int nz, int nghost,
                                 double dx, double dy, double dz, double gamma,
                                 const int bound[6], pgpu_domain **out);

/* USERDEF sides (boundary.c:258-292): user code on the host sets the ghost
   zones, before every stage.  The slab of the side is handed to the hook:
   the ghost zones of the side and the first nghost interior layers next to
   them, over the whole box (ghost zones and the plane below them included)
   along the other two axes, with the current values of the stage input.
   Indices are those of PLUTO's ghosted box (interior cell (0,0,0) of the
   native API is (nghost, nghost, nghost); -1 is the plane below the box,
   which holds the face below the first cells for the staggered arrays and
   nothing for the cell arrays).  Element (k, j, i) of an array is
   vc[q][((k - lo[2])*ld[1] + (j - lo[1]))*ld[0] + (i - lo[0])], the same
   for vs[a] (the face above cell (k, j, i) along axis a).  The hook writes
   the ghost zones it owns (and may write anything in the slab); the whole
   slab is copied back.  Then, as boundary.c:285-290, the device fills the
   normal ghost faces from div.B = 0 (FillMagneticField) and the
   cell-centred normal field from them (CT_AverageNormalMagField); the
   cell-centred transverse field (CT_AverageTransverseMagField) is the
   hook's to set.  side: 0..5 = X1_BEG, X1_END, ... X3_END; stage: 1 or 2
   (g_intStage).  A nonzero return stops the step with PGPU_EINVAL.  The
   hook runs on the calling thread inside pgpu_run, which then synchronises
   at every stage. */
typedef struct pgpu_slab {
  int lo[3], hi[3];   /* index ranges (i, j, k), inclusive */
  int ld[3];          /* hi - lo + 1 */
  int beg[3], end[3]; /* the interior cells of the box (IBEG..IEND, ...) */
  int tot[3];         /* cells per axis, ghosts included (NX1_TOT, ...) */
  double *vc[8];      /* rho, vx, vy, vz, Bx, By, Bz, p */
  double *vs[3];      /* b_x, b_y, b_z faces */
} pgpu_slab;
typedef int (*pgpu_userdef_fn) (void *ctx, int side, int stage, pgpu_slab *slab);
int  pgpu_set_userdef_boundary (pgpu_domain *dom, pgpu_userdef_fn fn, void *ctx);

/* Branch signature (a diagnostic with no PLUTO counterpart): while on, every
   stage of pgpu_run also recomputes each comparison whose outcome selects an
   expression of the step (the MC limiter cases, the min/max of the HLLD
   wave-speed estimates and its regions and degeneracy switch, the UCT_HLL
   face-speed maxima and MC_LIM2 slopes) and adds a 64-bit hash of each
   cell's outcomes into one word; ConsToPrim floors count too.  Runs from
   nearby initial states with equal signatures followed the same smooth
   piece of the step, so their finite differences are derivatives of that
   piece.  Periodic fused FP64 HLLD + UCT_HLL only (PGPU_EINVAL otherwise).
   Turning it on clears the word; the extra kernels cost about one flux
   sweep per stage. */
int  pgpu_set_branch_signature (pgpu_domain *dom, int on);
int  pgpu_get_branch_signature (pgpu_domain *dom, unsigned long long *sig);

/* Branch tape (a diagnostic with no PLUTO counterpart).  PGPU_TAPE_RECORD:
   the steps that follow run the modular kernels (bit-identical to the fused
   ones) and also store, per cell and stage, the outcome of every comparison
   listed above (32 bytes per cell per stage; capacity max_steps steps,
   allocated here).  PGPU_TAPE_REPLAY: the steps that follow take those
   outcomes from the tape, from its first step on, instead of comparing, so
   they evaluate, at whatever state they start from, the smooth piece of
   the step that the recorded run lay on; the central differences of a
   replayed objective are the derivatives the discrete adjoint computes.
   Running past the recorded steps returns PGPU_EINVAL.  ConsToPrim's floors
   are not taped.  PGPU_TAPE_OFF frees the tape.  Periodic FP64 HLLD +
   UCT_HLL only; not while the branch signature is on (PGPU_EINVAL). */
#define PGPU_TAPE_OFF     0
#define PGPU_TAPE_RECORD  1
#define PGPU_TAPE_REPLAY  2
int  pgpu_set_branch_tape (pgpu_domain *dom, int mode, int max_steps);

/* Select the flux-path precision (PGPU_PREC_*).  PGPU_PREC_FP32_FLUX needs
   the fused pipeline: PLM and HLLD run in FP32 on states rounded from FP64;
   the flux differences, the conservative accumulation, the EMF, the CT
   update and ConsToPrim stay FP64. */
int  pgpu_set_precision (pgpu_domain *dom, int precision);

/* Select the kernel pipeline (PGPU_PIPELINE_*). */
int  pgpu_set_pipeline (pgpu_domain *dom, int pipeline);

/* Select the Riemann solver of the sweeps (PGPU_SOLVER_*), as PLUTO's
   SetSolver (set_solver.c:13-81). */
int  pgpu_set_solver (pgpu_domain *dom, int solver);

/* EMF average of the constrained-transport update: PLUTO's CT_EMF_AVERAGE
   (ct.h:18-21; a compile-time macro there, chosen here per domain) with
   its label values.  UCT_HLL is the default and the only one of the FP32
   flux path, the modular pipeline and the adjoint; the other three run on
   the FP64 fused pipeline with any of the solvers above.
     ARITHMETIC   CT_EMF_ArithmeticAverage(emf, 0.25), ct_emf.c:238-241
     UCT0         face EMFs corrected by the cell EMF, ct_emf.c:261-283
     UCT_CONTACT  CT_EMF_IntegrateToCorner, ct_emf.c:242-255,
                  ct_emf_average.c:46-177
     UCT_HLL      CT_EMF_HLL_Solver, ct_emf.c:256-260 */
#define PGPU_EMF_ARITHMETIC  1
#define PGPU_EMF_UCT0        2
#define PGPU_EMF_UCT_CONTACT 3
#define PGPU_EMF_UCT_HLL     4
int  pgpu_set_emf_average (pgpu_domain *dom, int emf_average);

/* Tiling of the z-marching UCT_HLL EMF kernel, process-wide: workgroups of
   bx x by cells (bx*by a multiple of 64, at most 256) marching up to kz
   planes (halved while the box would get fewer than 2048 workgroups);
   (0, 0, 0) restores the default 64 x 4 x 16.  A tuning knob with no PLUTO
   counterpart; results do not depend on it. */
int  pgpu_set_emf_tiling (int bx, int by, int kz);
int  pgpu_get_emf_tiling (int *bx, int *by, int *kz);
/* UCT_HLL EMF kernel of FP64 3D boxes (process-wide; results identical):
   0 the register-rolling z-march (tiled by pgpu_set_emf_tiling); the
   LDS-staged z-march (boxes with nx even) with column tiles of 1: 64 x 4,
   2: 32 x 4, 3: 32 x 8 (the default); kz planes per workgroup (0: 16). */
int  pgpu_set_emf_kernel (int kind, int kz);
/* x/y flux kernel of the fused step (process-wide; FP64 results identical,
   FP32-flux results equal to FP32 rounding, 5e-7 relative):
   0 overlapping z-plane tiles; 1 (the default) one workgroup per box row
   marching up y segments of seg rows (0: 32, fewer on boxes too small to
   fill the GPU with 32-row segments), each cell reconstructed
   once per direction, on boxes with nx even (a multiple of 4 for the
   FP32-flux path) and <= 256, ny >= 4 and fields under 4 GiB; the tiles
   elsewhere. */
int  pgpu_set_flux_kernel (int kind, int seg);

/* Host <-> device state copies (synchronous).  prim is [8][nz][ny][nx];
   the cell-centred field in prim is ignored on input and recomputed as the
   face average 0.5*(b_{i+1/2} + b_{i-1/2}) (startup.c:235-241). */
int  pgpu_set_state (pgpu_domain *dom, const double *prim, const double *bx,
                     const double *by, const double *bz);
int  pgpu_get_state (pgpu_domain *dom, double *prim, double *bx, double *by,
                     double *bz);

/* pgpu_set_state of a bounded box with the box's own lower boundary faces,
   as PLUTO's Startup sets them (startup.c:177-204; Vs[BX1s] at IBEG-1 and
   its like): bx_low[nz][ny] is b_x at the face below cell i = 0,
   by_low[nz][nx] b_y below j = 0, bz_low[ny][nx] b_z below k = 0.  CT
   evolves these faces (ct_update.c ranges start at IBEG-1), so a
   non-periodic solenoidal field needs them for div.B = 0 in the first cell
   layer.  Refused on a periodic box. */
int  pgpu_set_state_bounded (pgpu_domain *dom, const double *prim, const double *bx,
                             const double *by, const double *bz, const double *bx_low,
                             const double *by_low, const double *bz_low);

/* pgpu_get_state of a bounded box, with the faces below the first cells in
   the layout of pgpu_set_state_bounded.  Refused on a periodic box. */
int  pgpu_get_state_bounded (pgpu_domain *dom, double *prim, double *bx, double *by,
                             double *bz, double *bx_low, double *by_low, double *bz_low);

/* Device pointers of the resident state (prim[8], face[3]), for callers
   that exchange device buffers directly.  The face buffers rotate: the
   pointers are valid until the next pgpu_run. */
int  pgpu_state_device_ptrs (pgpu_domain *dom, double **prim8, double **face3);

/* Time-step control.  dt is the next step's dt (PLUTO's first_dt at start);
   with fixed_dt != 0 every step uses dt, otherwise after each step
   dt = min(cfl/invDt_hyp, cfl_max_var*dt) as in NextTimeStep. */
int  pgpu_set_dt (pgpu_domain *dom, double dt, double cfl, double cfl_max_var,
                  int fixed_dt);

/* Enqueue nsteps RK2 steps on the domain's stream (asynchronous). */
int  pgpu_run (pgpu_domain *dom, int nsteps);
/* Waits for the domain's stream.  Returns PGPU_ENAN once a stage input has
   held a NaN (the state set by pgpu_set_state, or the primitive state a
   stage produced): PLUTO's CheckNaN (tools.c:20-38, called on every row of
   every stage, update_stage.c:192) stops the run at that point.  Returns
   PGPU_EDTSMALL once the device dt control (not fixed_dt) computed a next
   dt below 1e-9 of the first dt given to pgpu_set_dt: PLUTO's NextTimeStep
   (main.c:534-543) stops the run there; the steps enqueued after it take
   dt = 0 and leave the time unchanged. */
int  pgpu_sync (pgpu_domain *dom);

/* Scalars after the last completed step (synchronises):
   out[0] next dt, out[1] invDt_hyp of the last step, out[2] its max Mach,
   out[3] time, out[4] steps taken, out[5] ConsToPrim floor events. */
int  pgpu_get_scalars (pgpu_domain *dom, double out[6]);

/* dt, invDt_hyp and max Mach of steps [0, n) since the last pgpu_set_dt
   (n <= 4096). */
int  pgpu_get_history (pgpu_domain *dom, int n, double *dt, double *invdt,
                       double *mach);

/* ConsToPrim failure flags of the last step (bit 64 = FLAG_CONS2PRIM_FAIL). */
int  pgpu_get_flags (pgpu_domain *dom, unsigned char *flag);

/* Kernel timing with HIP events recorded on the domain's stream around
   every kernel launch.  Kernel ids: 0 sweep-x, 1 sweep-y, 2 sweep-z,
   3 update, 4 emf, 5 ct (ct+c2p in the fused pipeline), 6 c2p, 7 step-end,
   8 flux-xy, 9 flux-z, 10 boundary (the ghost fill of a bounded box, all
   six sides).  read: total milliseconds and
   launch count per id since enable (synchronises). */
int  pgpu_timing_enable (pgpu_domain *dom, int enable);
int  pgpu_timing_read (pgpu_domain *dom, int nid, double *ms, long *count);
int  pgpu_kernel_name (int id, const char **name);

/* Wall time of nsteps steps bracketed by events on the domain's stream. */
int  pgpu_time_steps (pgpu_domain *dom, int nsteps, double *ms);

/* Bytes of device memory held by the domain. */
int  pgpu_domain_bytes (pgpu_domain *dom, size_t *bytes);

/* Per-field difference of dom against ref (same shape, same device, e.g.
   the FP32-flux path against the FP64 path).  Fields f = 0..7: rho, m_x,
   m_y, m_z, E (from the cell-centred primitives), b_x, b_y, b_z (faces).
   out[4f+0] = sum|dom-ref|, out[4f+1] = sum|ref|, out[4f+2] = max|dom-ref|,
   out[4f+3] = max|ref|.  Synchronises both domains' streams. */
int  pgpu_state_error (pgpu_domain *ref, pgpu_domain *dom, double out[32]);

/* ------------------------------------------------------------------ */
/* PLUTO's .dbl checkpoint format (DBL_OUTPUT single_file + restart)    */
/* ------------------------------------------------------------------ */

/* Write output #nfile of a periodic box in PLUTO's double-precision format
   into dir: data.NNNN.dbl (write_data.c:111-170 with FileWriteData,
   bin_io.c:216-259: the cell-centred variables, then the staggered
   fields with the image face below the first cell), the dbl.out line
   (write_data.c:365-395) and the restart.out record (restart.c:260-299),
   so PLUTO can restart from them and read them like its own files.
   Variables: DIMENSIONS = COMPONENTS = 3 or 2 (nz = 1) MHD with CT.
   Layouts of prim/bx/by/bz as pgpu_set_state; time, dt (the next step's)
   and nstep are written as g_time, g_dt, g_stepNumber.  The _host forms
   run on the host only (no device needed). */
int  pgpu_write_dbl_host (const char *dir, int nfile, int nx, int ny, int nz, int dims,
                          const double *prim, const double *bx, const double *by,
                          const double *bz, double time, double dt, long nstep);
/* Read output #nfile (RestartFromFile, restart.c:17-257): the dbl.out line
   gives the endianness (big-endian files are byte-swapped), restart.out
   time, dt and nstep, the data file the state (image faces dropped;
   prim's vx3, Bx3 and bz are zero for 2D files). */
int  pgpu_read_dbl_host (const char *dir, int nfile, int nx, int ny, int nz, int dims,
                         double *prim, double *bx, double *by, double *bz, double *time,
                         double *dt, long *nstep);
/* The same for a domain: writes its state with its time, next dt and step
   count; reading sets its state and restores dt, time and step (the CFL
   settings of the last pgpu_set_dt stay). */
int  pgpu_write_dbl (pgpu_domain *dom, const char *dir, int nfile);
int  pgpu_read_dbl (pgpu_domain *dom, const char *dir, int nfile);

/* ------------------------------------------------------------------ */
/* Device-resident state exchange and the discrete adjoint             */
/* ------------------------------------------------------------------ */

/* State copies from/to device buffers (asynchronous on the domain's
   stream): prim8 is [8][nz][ny][nx], face3 is [3][nz][ny][nx].  On input
   the cell-centred field in prim8 is ignored and recomputed as the face
   average. */
int  pgpu_set_state_device (pgpu_domain *dom, const double *prim8, const double *face3);
int  pgpu_get_state_device (pgpu_domain *dom, double *prim8, double *face3);

/* Discrete adjoint of one RK2 step (FP64, DIMENSIONS 3, fixed dt set with
   pgpu_set_dt(..., fixed_dt = 1)), for gradients of objectives of the final
   state.  The domain holds the step's input state (V^n, b^n); lam_prim
   ([8][n], device) and lam_face ([3][n], device) hold on input the adjoint
   of the step's output, dJ/dV^{n+1} and dJ/db^{n+1} (the cell-centred field
   components of lam_prim count as the face average's adjoint), and on
   output the adjoint of the input: dJ/dV^n with its field components zero
   (folded into the faces) and dJ/db^n.  The step is recomputed internally;
   the domain's state is left unchanged.  Non-smooth points take the branch
   the forward step took (MC limiter, HLLD region, min/max); ConsToPrim
   floors are assumed inactive.  Synchronise with pgpu_sync. */
int  pgpu_adjoint_step (pgpu_domain *dom, double *lam_prim, double *lam_face);

/* The same adjoint with the forward step stored instead of recomputed (the
   device memory of a run's steps traded for a forward step per adjoint
   step).  pgpu_adjoint_record runs one step from the domain's state, bit for
   bit the step pgpu_run takes with the fixed dt (the step counter, time and
   dt history are not advanced and no NaN check is made), writes what the
   reverse of that step reads into tape (device, PGPU_ADJOINT_TAPE_ARRAYS
   arrays of nx*ny*nz doubles: the input state, the stage-1 output, both
   stages' face speeds and velocity slopes, the output primitives) and
   leaves the domain at the step's output.  pgpu_adjoint_step_taped is
   pgpu_adjoint_step for the step a tape holds, bit-identical to it; it does
   not read the domain's state.  Same requirements as pgpu_adjoint_step. */
#define PGPU_ADJOINT_TAPE_ARRAYS 60
int  pgpu_adjoint_record (pgpu_domain *dom, double *tape);
int  pgpu_adjoint_step_taped (pgpu_domain *dom, const double *tape, double *lam_prim,
                              double *lam_face);

/* Point-level tangent and adjoint of HLLD for checking the adjoint (host
   buffers).  Rows are [n][8] primitive states (normal field slot = the
   face field) along direction dir (0, 1, 2).  Outputs per row, 8 values:
   fluxes of rho, m_x, m_y, m_z, E, the total pressure, SL, SR.
   tangent: out = HLLD(vL, vR), dout = J (tL, tR);
   adjoint: aL, aR = J^T lam. */
int  pgpu_hlld_tangent_rows (int dir, int n, double gamma, const double *vL,
                             const double *vR, const double *tL, const double *tR,
                             double *out, double *dout);
int  pgpu_hlld_adjoint_rows (int dir, int n, double gamma, const double *vL,
                             const double *vR, const double *lam, double *aL,
                             double *aR);

/* ------------------------------------------------------------------ */
/* Batched patches                                                     */
/* ------------------------------------------------------------------ */

/* A batch holds npatch independent periodic cubes (AMR-like patches) of
   edge[p] cells (3 <= edge <= 1024) on the unit box (dx = 1/edge), FP64,
   DIMENSIONS 3.  Every step advances all patches with one persistent
   launch per RK stage (workgroups take whole patches from a device work
   queue, largest first), both launches replayed from one hipGraph; each
   patch follows its own CFL time step.  A patch's result is bit for bit
   that of a pgpu_domain of the same size. */
typedef struct pgpu_batch pgpu_batch;

int  pgpu_batch_create (int npatch, const int *edge, double gamma, pgpu_batch **out);
int  pgpu_batch_destroy (pgpu_batch *b);
/* Patch state copies (synchronous; layouts as pgpu_set_state/get_state). */
int  pgpu_batch_set_patch (pgpu_batch *b, int p, const double *prim, const double *bx,
                           const double *by, const double *bz);
int  pgpu_batch_get_patch (pgpu_batch *b, int p, double *prim, double *bx, double *by,
                           double *bz);
/* Every patch restarts from dt; then per patch as pgpu_set_dt. */
int  pgpu_batch_set_dt (pgpu_batch *b, double dt, double cfl, double cfl_max_var, int fixed_dt);
int  pgpu_batch_run (pgpu_batch *b, int nsteps);
/* Waits for the batch; PGPU_EDTSMALL once a patch's dt control reached
   NextTimeStep's stop (main.c:534-543; that patch then takes dt = 0). */
int  pgpu_batch_sync (pgpu_batch *b);
int  pgpu_batch_time_steps (pgpu_batch *b, int nsteps, double *ms);
/* Patch p after its last step: out[0] next dt, [1] invDt_hyp, [2] max Mach,
   [3] time, [4] steps, [5] ConsToPrim floor events. */
int  pgpu_batch_patch_scalars (pgpu_batch *b, int p, double out[6]);
/* Launch shape: workgroup size (64, 128, 256, 512, 1024) and number of
   resident workgroups (<= 0: 4 per CU). */
int  pgpu_batch_set_launch (pgpu_batch *b, int threads, int groups);
int  pgpu_batch_get_launch (pgpu_batch *b, int *threads, int *groups);
/* The cells a workgroup takes in one pass over a phase of a patch: a
   tx x ty x tz box (tx*ty*tz = the workgroup size; the boxes in storage
   order, so the stencil neighbours of a box's cells are read close in
   time), or (0, 0, 0): consecutive cells in storage order (the default).
   Results do not depend on it.  A launch change that no longer fits the
   tile returns to the linear order. */
int  pgpu_batch_set_tile (pgpu_batch *b, int tx, int ty, int tz);
/* How the step runs (results identical): 1 (the default) stacks the patches
   whose edge is even and 4 to 256 cells by edge and runs each stack through
   the fused kernels of a single domain, z periodic within each patch and
   every patch with its own scalars; the other patches, and all of them in
   mode 0, go through the persistent patch kernel.  pgpu_batch_get_mode
   returns the mode and how many patches the captured step stacks. */
int  pgpu_batch_set_mode (pgpu_batch *b, int mode);
int  pgpu_batch_get_mode (pgpu_batch *b, int *mode, int *stacked_patches);
int  pgpu_batch_get_tile (pgpu_batch *b, int *tx, int *ty, int *tz);
/* Autotuner: times workgroup sizes {128, 256, 512, 1024} x {1, 2, 4, 8}
   workgroups per CU with the linear cell order, then every cell tile of the
   fastest workgroup size (tx a power of two from 4 to 64, ty >= tz), over
   trial_steps steps each (the trial steps advance the patches), and keeps
   the fastest launch shape and tile (pgpu_batch_get_tile). */
int  pgpu_batch_tune (pgpu_batch *b, int trial_steps, int *threads, int *groups,
                      double *ms_per_step);
int  pgpu_batch_bytes (pgpu_batch *b, size_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
