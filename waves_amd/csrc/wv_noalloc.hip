// Allocation-free liquid volume: the FP64 fixed-order volume reduction of
// wv_liquid_volume over a caller's device volume-fraction field, with its
// per-group partials in a file-scope device array instead of solver-owned
// memory.  For callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The array `scratch` is a host-side stand-in in host code: its device address
// is taken once per device with hipGetSymbolAddress (each GPU holds its own
// copy) and only that address is handed to kernels.  Every call on one GPU
// shares that GPU's array: calls that overlap on one GPU race.  Each call
// rewrites every partial it reads, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>
#include <mutex>
#include "wv_device.h"
#include "../../include/waves.h"

#define NOALLOC_GROUPS (WV_NOALLOC_MAX_CELLS / WV_RED)
#define NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[NOALLOC_GROUPS];

namespace {
std::mutex g_lock;
double* g_scratch[NOALLOC_MAX_DEVICES] = {nullptr};

// the device address of `scratch` on the current device, looked up once per device
int scratch_address(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return WV_ENODEV;
  std::lock_guard<std::mutex> g(g_lock);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return WV_EHIP;
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return WV_OK;
}
}  // namespace

// The 128-lane FP64 group sum of the solver's reductions (an LDS exchange of
// the upper 64 lanes, then shuffles), on the x component alone: the same
// additions in the same order as the volume's four-component form.
__device__ __forceinline__ double group_sum_d(double v) {
  __shared__ double upper[64];
  const int lid = threadIdx.x;
  if (lid >= 64) upper[lid - 64] = v;
  __syncthreads();
  if (lid < 64) {
    v += upper[lid];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  }
  __syncthreads();
  return v;
}
__global__ void __launch_bounds__(WV_RED) k_volume_partials(const float* __restrict__ F, int64_t n, double* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * WV_RED + threadIdx.x;
  const double v = group_sum_d(i < n ? (double)F[i] : 0.0);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}
__global__ void __launch_bounds__(WV_RED) k_volume_final(const double* __restrict__ part, int64_t groups, double cell, double* __restrict__ out) {
  double v = 0;
#pragma unroll 8
  for (int64_t k = threadIdx.x; k < groups; k += WV_RED) v += part[k];
  v = group_sum_d(v);
  if (threadIdx.x == 0) *out = v * cell;
}

extern "C" int wv_liquid_volume_noalloc(const float* fraction, int64_t cells, float hx, float hy, float hz, double* volume, void* stream) {
  if (!fraction || !volume || cells < 1) return WV_EINVAL;
  if (cells > WV_NOALLOC_MAX_CELLS) return WV_EINVAL;   // beyond the static scratch
  double* part = nullptr;
  int rc = scratch_address(&part);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t groups = (cells + WV_RED - 1) / WV_RED;
  hipLaunchKernelGGL(k_volume_partials, dim3((unsigned)groups), dim3(WV_RED), 0, st, fraction, cells, part);
  hipLaunchKernelGGL(k_volume_final, dim3(1), dim3(WV_RED), 0, st, (const double*)part, groups,
                     (double)hx * (double)hy * (double)hz, volume);
  return hipGetLastError() == hipSuccess ? WV_OK : WV_EHIP;
}
