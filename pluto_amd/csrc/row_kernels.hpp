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

}  // namespace pgpu
