// Launchers of the row-granular kernels (row_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace pgpu {

// solver: SOLVER_HLLD, _HLL, _HLLC, _TVDLF or _ROE (mhd_device.hpp); full
// fluxes (B included)
int launch_riemann_rows(int solver, int dir, int n, const double* vl, const double* vr,
                        double gamma, double* flux, double* press, double* sl, double* sr,
                        double* cmax, double* mach, hipStream_t s);
int launch_plm_rows(int dir, int n, const double* v, const double* bn, double gamma,
                    double* vp, double* vm, double* up, double* um, hipStream_t s);
int launch_p2c_rows(int n, const double* v, double gamma, double* u, hipStream_t s);
int launch_c2p_rows(int n, double* u, double gamma, double* v, unsigned char* flag,
                    hipStream_t s);
int launch_face_average_box(int mx, int my, int mz, const double* fx, const double* fy,
                            const double* fz, int dims, double* out, hipStream_t s);

// CT on PLUTO's ghosted boxes (n[0] x n[1] x n[2] cells, [k][j][i] order), over
// the EMF range lo..hi (inclusive; 1 <= lo, hi <= n - 3 along each axis):
// the UCT_HLL edge EMFs from the 21 fields of EmfField order in f21 (each a
// box), written to ex, ey, ez; and the face update of CT_Update from ex, ey,
// ez, the faces b0 (x, y, z) and the cell widths dx1, dx2, dx3, written to b.
int launch_ct_emf_box(const int n[3], const int lo[3], const int hi[3], const double* const* f21,
                      double* ex, double* ey, double* ez, hipStream_t s);
int launch_ct_update_box(const int n[3], const int lo[3], const int hi[3], double dt,
                         const double* dx1, const double* dx2, const double* dx3,
                         const double* ex, const double* ey, const double* ez,
                         const double* const* b0, double* const* b, hipStream_t s);

}  // namespace pgpu
