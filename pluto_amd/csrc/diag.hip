// Diagnostics of resident domains: per-field difference norms between two
// domains of the same shape (used to state the FP32-flux path's error against
// the FP64 path on the device, without copying whole states to the host).

#include <hip/hip_runtime.h>

#include "domain.hpp"
#include "mhd_device.hpp"

namespace pgpu {

namespace {

constexpr int kFields = 8;          // rho, m_x, m_y, m_z, E, b_x, b_y, b_z
constexpr int kStats = 4 * kFields; // sum|d|, sum|ref|, max|d|, max|ref|
constexpr int kBlock = 256;
constexpr int kMaxBlocks = 2048;

struct Pair {
  const double* a[8 + 3];   // prim 8, faces 3
  const double* b[8 + 3];
};

__device__ __forceinline__ void fields_of(const double* const* p, long c, double gmm1, double* f)
{
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) v[q] = p[q][c];
  f[0] = v[RHO];
  f[1] = v[RHO] * v[VX1];
  f[2] = v[RHO] * v[VX2];
  f[3] = v[RHO] * v[VX3];
  f[4] = total_energy(v, gmm1);
  f[5] = p[8][c];
  f[6] = p[9][c];
  f[7] = p[10][c];
}

// Per-block partial statistics (fixed order within a block).
__global__ void __launch_bounds__(kBlock) k_state_error(long n, double gmm1, Pair pr, double* part)
{
  __shared__ double red[kBlock];
  double st[kStats];
#pragma unroll
  for (int q = 0; q < kStats; q++) st[q] = 0.0;
  for (long c = (long)blockIdx.x * kBlock + threadIdx.x; c < n; c += (long)gridDim.x * kBlock) {
    double fa[kFields], fb[kFields];
    fields_of(pr.a, c, gmm1, fa);
    fields_of(pr.b, c, gmm1, fb);
#pragma unroll
    for (int f = 0; f < kFields; f++) {
      const double d = fabs(fb[f] - fa[f]), r = fabs(fa[f]);
      st[4 * f + 0] += d;
      st[4 * f + 1] += r;
      st[4 * f + 2] = d > st[4 * f + 2] ? d : st[4 * f + 2];
      st[4 * f + 3] = r > st[4 * f + 3] ? r : st[4 * f + 3];
    }
  }
  for (int q = 0; q < kStats; q++) {
    red[threadIdx.x] = st[q];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        const double x = red[threadIdx.x], y = red[threadIdx.x + w];
        red[threadIdx.x] = (q % 4) < 2 ? x + y : (y > x ? y : x);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.x * kStats + q] = red[0];
    __syncthreads();
  }
}

__global__ void k_state_error_final(int nblk, const double* part, double* out)
{
  const int q = threadIdx.x;
  if (q >= kStats) return;
  double acc = 0.0;
  for (int b = 0; b < nblk; b++) {
    const double x = part[(long)b * kStats + q];
    acc = (q % 4) < 2 ? acc + x : (x > acc ? x : acc);
  }
  out[q] = acc;
}

}  // namespace

}  // namespace pgpu

using namespace pgpu;

extern "C" int pgpu_state_error(pgpu_domain* ref, pgpu_domain* dom, double out[32])
{
  if (!ref || !dom || !out) return fail(PGPU_EINVAL, "NULL argument");
  if (ref->g.nx != dom->g.nx || ref->g.ny != dom->g.ny || ref->g.nz != dom->g.nz)
    return fail(PGPU_EINVAL, "domains differ in shape");
  if (ref->device != dom->device) return fail(PGPU_EINVAL, "domains are on different devices");
  if (ref->g.ghost || dom->g.ghost) return fail(PGPU_EINVAL, "the state error is for periodic boxes");
  if (hipStreamSynchronize(dom->stream) != hipSuccess)
    return fail(PGPU_EDEVICE, "stream synchronisation failed");
  const long n = ref->g.n;
  const int nblk = (int)((n + kBlock - 1) / kBlock < kMaxBlocks ? (n + kBlock - 1) / kBlock
                                                                 : kMaxBlocks);
  double* buf = nullptr;
  const size_t bytes = ((size_t)nblk + 1) * kStats * sizeof(double);
  if (hipMalloc(&buf, bytes) != hipSuccess) return fail(PGPU_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
  Pair pr;
  for (int q = 0; q < 8; q++) {
    pr.a[q] = ref->V[q];
    pr.b[q] = dom->V[q];
  }
  for (int a = 0; a < 3; a++) {
    pr.a[8 + a] = ref->bA[a];
    pr.b[8 + a] = dom->bA[a];
  }
  double* part = buf + kStats;
  hipLaunchKernelGGL(k_state_error, dim3(nblk), dim3(kBlock), 0, ref->stream, n, ref->g.gmm1, pr,
                     part);
  hipLaunchKernelGGL(k_state_error_final, dim3(1), dim3(64), 0, ref->stream, nblk, part, buf);
  hipError_t err = hipGetLastError();
  if (err == hipSuccess)
    err = hipMemcpyAsync(out, buf, kStats * sizeof(double), hipMemcpyDeviceToHost, ref->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(ref->stream);
  (void)hipFree(buf);
  if (err != hipSuccess) return fail(PGPU_EDEVICE, "state error kernel: %s", hipGetErrorString(err));
  return PGPU_OK;
}
