// The FP32-flux instantiations of the fused flux kernels (fused_flux.hpp).
// Their PLM and Riemann arithmetic is FP32 and already differs from the
// reference at the 1e-7 level, so this translation unit lets the compiler
// contract multiply-add pairs into FMAs (everywhere else the library is
// compiled without contraction, which keeps the FP64 path bit-exact).  The
// pragma precedes the headers, so it covers the device templates they
// define as instantiated here.

#include <hip/hip_runtime.h>

#pragma clang fp contract(fast)

#include "fused_flux.hpp"

namespace pgpu {

int launch_flux_fused_f32(const Geo& g, int stage, const StageIn& in, const Cons& U,
                          const FaceData& fd, double* cdt_xy, DevScalars* sc, hipStream_t s,
                          KernelTimer* kt)
{
  if (g.emf_avg != EMF_UCT_HLL) return (int)hipErrorInvalidValue;   // UCT_HLL face data only
  if (g.solver == SOLVER_HLL) launch_flux_fused_t<float, SOLVER_HLL>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  else if (g.solver == SOLVER_HLLC) launch_flux_fused_t<float, SOLVER_HLLC>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  else if (g.solver == SOLVER_TVDLF) launch_flux_fused_t<float, SOLVER_TVDLF>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  else if (g.solver == SOLVER_ROE) launch_flux_fused_t<float, SOLVER_ROE>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  else launch_flux_fused_t<float, SOLVER_HLLD>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  return (int)hipGetLastError();
}

}  // namespace pgpu
