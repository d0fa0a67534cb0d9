// Tolerance-mode FP64 flux kernels (PGPU_PREC_FP64_FAST): the FP64
// instantiation of the fused flux kernels (fused_flux.hpp) compiled with FMA
// contraction and with Newton-refined division and square root
// (PGPU_FAST_F64 in mhd_device.hpp).  Results are no longer bit-exact with
// the reference; they agree to ~1e-15 per operation, within the 1e-12
// relative per step that the tolerance mode is held to.  The kernels carry
// V = 1, so their symbols differ from the bit-exact FP64 build's.

#include <hip/hip_runtime.h>

#pragma clang fp contract(fast)
#define PGPU_FAST_F64 1

#include "fused_flux.hpp"

namespace pgpu {

int launch_flux_fused_fast(const Geo& g, int stage, const StageIn& in, const Cons& U,
                           const FaceData& fd, double* cdt_xy, DevScalars* sc, hipStream_t s,
                           KernelTimer* kt)
{
  if (g.emf_avg != EMF_UCT_HLL || g.solver != SOLVER_HLLD) return (int)hipErrorInvalidValue;
  if (g.pz > 0 && (g.dims != 3 || !flux_rows_ok(g, 8, g.solver))) return (int)hipErrorInvalidValue;
  if (g.dims == 3 && g.n * 8 >= (1l << 32)) return (int)hipErrorInvalidValue;   // z march: 32-bit offsets
  launch_flux_fused_t<double, SOLVER_HLLD, FACE_UCT_HLL, 1>(g, stage, in, U, fd, cdt_xy, sc, s, kt);
  return (int)hipGetLastError();
}

}  // namespace pgpu
