// Fused stage kernels (fused_kernels.hip): tile shapes and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include "stage_kernels.hpp"

namespace pgpu {

constexpr int XY_TX = 32, XY_TY = 8;   // k_flux_xy workgroup (one z-plane); FP64 tiles overlap
constexpr int Z_TX = 64, Z_TY = 4;     // k_flux_z columns per workgroup

// Per-face UCT_HLL speeds and per-cell velocity slopes written by the flux
// kernels for the EMF kernel (same arrays as SweepOut's).
struct FaceData {
  double* SL[3];
  double* SR[3];
  double* dv[3][3];   // dv[direction][velocity component]
};

// fp32_flux: PLM and HLLD in FP32 (states rounded from FP64 on load); the
// flux differences and U are FP64; the face speeds and velocity slopes are
// stored as float in the first half of fd's buffers (launch_emf with
// f32_faces reads them).
int launch_flux_fused(const Geo& g, int stage, bool fp32_flux, const StageIn& in, const Cons& U,
                      const FaceData& fd, double* cdt_xy, DevScalars* sc, hipStream_t s,
                      KernelTimer* kt);
// the tolerance-mode FP64 launches (fused_fast.hip: FMA contraction, Newton
// division and square root; HLLD, UCT_HLL)
int launch_flux_fused_fast(const Geo& g, int stage, const StageIn& in, const Cons& U,
                           const FaceData& fd, double* cdt_xy, DevScalars* sc, hipStream_t s,
                           KernelTimer* kt);
// x/y flux kernel of the fused stage (process-wide; FP64 results identical,
// FP32-flux ones to FP32 rounding: the two kernels contract differently):
// 0 the overlapping z-plane tiles (k_flux_xy, k_flux_xy_items), 1 whole
// rows marching up y segments of seg rows (k_flux_xy_rows; 0: by the box's
// size, flux_xy_seg) on boxes with nx even (FP32 flux: a multiple of 4),
// 4 <= nx <= 256, ny >= 4 and fields under 4 GiB; the tiles elsewhere
int set_flux_xy_kernel(int kind, int seg);
int flux_xy_kind();
int flux_xy_seg(const Geo& g);
// the row kernel takes this box (flux real type of real_bytes, solver)
inline bool flux_rows_ok(const Geo& g, int real_bytes, int solver)
{
  const int nt = (g.nx + 63) / 64 * 64;
  (void)solver;
  return flux_xy_kind() == 1 && g.nx >= 4 &&
         g.nx % (real_bytes == 4 ? 4 : 2) == 0 && nt <= 256 && g.ny >= 4 &&
         g.n * 8 < (1l << 32);
}
// the FP32-flux launches (fused_f32.hip, compiled with FMA contraction)
int launch_flux_fused_f32(const Geo& g, int stage, const StageIn& in, const Cons& U,
                          const FaceData& fd, double* cdt_xy, DevScalars* sc, hipStream_t s,
                          KernelTimer* kt);

// CT face update fused with the face average and ConsToPrim.  Stage 2
// blends with the step-start faces b0 and writes the new faces to bout,
// which must not alias bin or b0.
// f32_out (stage 1 of the FP32-flux path): Vout, the next stage's input,
// is stored as float in the first half of each buffer; the stage-2 flux and
// EMF kernels of that path read it so.
int launch_ct_c2p(const Geo& g, int stage, const EdgeOut& e, const FaceIn& bin, const FaceIn& b0,
                  const FaceOut& bout, const Cons& U, const StageIn& vn, const Prims& Vout,
                  unsigned char* flag, DevScalars* sc, hipStream_t s, KernelTimer* kt,
                  bool f32_out = false);

}  // namespace pgpu
