// Standalone pressure solve and its mixed-precision form.
//
// Plain form: the reference's PCG op sequence (pressure_iterations.glslinc:4-32
// with the V-cycle hooks of water/pressure_gpu.gd:73-74) on the current
// geometry, from a zero pressure.  Mixed form: the same iteration with the
// V-cycle's operators stored in BF16 (FP32 vectors and accumulation), every
// CG dot product accumulated in FP64, and convergence confirmed by an FP64
// recomputation of the true residual b - A p; if the true residual misses the
// tolerance the recurrence residual is replaced by it and CG restarts from the
// current pressure.  The restart folds the FP32 pressure into an FP64
// accumulator and CG continues on the FP32 correction (mixed-precision
// iterative refinement), so the FP64 solution reaches tolerances below FP32's
// rounding floor.  No reference counterpart beyond the plain form.
#include "wv_mgdev.h"

static inline unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

struct d4 { double x, y, z, w; };
__device__ __forceinline__ d4 mkd(double x, double y, double z, double w) { d4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
// 128-lane group sum (x, y, z) and max (w) in FP64: LDS exchange then shuffles
__device__ __forceinline__ d4 group_reduce_d(d4 v) {
  __shared__ d4 upper[64];
  int lid = threadIdx.x;
  if (lid >= 64) upper[lid - 64] = v;
  __syncthreads();
  if (lid < 64) {
    d4 o = upper[lid]; v.x += o.x; v.y += o.y; v.z += o.z; v.w = v.w < o.w ? o.w : v.w;
    for (int off = 32; off > 0; off >>= 1) {
      double ox = __shfl_down(v.x, off, 64), oy = __shfl_down(v.y, off, 64), oz = __shfl_down(v.z, off, 64), ow = __shfl_down(v.w, off, 64);
      v.x += ox; v.y += oy; v.z += oz; v.w = v.w < ow ? ow : v.w;
    }
  }
  __syncthreads();
  return v;
}
__device__ __forceinline__ d4 global_reduce_d(const double4* partial, int64_t groups) {
  d4 v = mkd(0, 0, 0, 0);
  for (int64_t k = threadIdx.x; k < groups; k += WV_RED) { double4 p = partial[k]; v.x += p.x; v.y += p.y; v.z += p.z; v.w = v.w < p.w ? p.w : v.w; }
  return group_reduce_d(v);
}
__device__ __forceinline__ void put_partial(const Ctx& s, d4 v) { if (threadIdx.x == 0) s.partial64[blockIdx.x] = make_double4(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ bool converged64(const Ctx& s, double r2, double b2, double maxr) {
  double dt = timestep(s), tol = s.tol;
  double sb = dt * dt / 1000 * sqrt(b2 > 0 ? b2 : 0), sr = dt * dt / 1000 * sqrt(r2 > 0 ? r2 : 0);
  return (r2 <= tol * tol * b2 || (sb <= 1e-5 && sr <= 1e-8)) && dt * dt / 1000 * maxr < 5e-5;
}

// open a pressure trial on the current state (pressure_pass.glslinc:20-24 without the quota)
__global__ void k_solve_begin(Ctx s) {
  float* c = s.c; c[4] = 1; c[5] = 1; c[6] = 0; c[7] = 0; c[8] = 0; c[13] = 1; c[57] = 0; c[58] = 0;
  for (int k = 0; k < 16; k++) s.dctl[k] = 0;
}
__global__ void __launch_bounds__(256) k_solve_zero_p(Ctx s, int zero_pressure) {
  int64_t i = GID(); if (i >= s.N) return;
  if (zero_pressure) putv(s, candidate(s), P_(s) + i, 0);
  s.pacc[i] = 0;
}
void wv_solve_begin(const Ctx& s, bool zero_pressure, hipStream_t st) {
  hipLaunchKernelGGL(k_solve_begin, dim3(1), dim3(1), 0, st, s);
  hipLaunchKernelGGL(k_solve_zero_p, dim3(nb(s.N, 256)), dim3(256), 0, st, s, zero_pressure ? 1 : 0);
}

// zero-mean right-hand side over the wet cells (p = 0, so r = b, z = r / diag)
__global__ void __launch_bounds__(WV_RED) k_rhs_mean(Ctx s, int final) {
  if (!final) {
    int64_t i = GID(); d4 v = mkd(0, 0, 0, 0);
    if (i < s.N && wet(s, coord(i, dims(s)))) v = mkd(s.work[5 * s.N + i], 1, 0, 0);
    v = group_reduce_d(v); put_partial(s, v); return;
  }
  d4 g = global_reduce_d(s.partial64, (s.N + WV_RED - 1) / WV_RED);
  if (threadIdx.x == 0) s.dctl[8] = g.y > 0 ? g.x / g.y : 0;
}
__global__ void __launch_bounds__(256) k_rhs_shift(Ctx s) {
  int64_t i = GID(); int64_t N = s.N; if (i >= N) return;
  if (!wet(s, coord(i, dims(s)))) return;
  float b = (float)((double)s.work[5 * N + i] - s.dctl[8]);
  float z = b / s.work[12 * N + i];
  s.work[5 * N + i] = b; s.work[i] = b; s.work[N + i] = z; s.work[2 * N + i] = z;
}
void wv_solve_zero_mean(const Ctx& s, hipStream_t st) {
  hipLaunchKernelGGL(k_rhs_mean, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0);
  hipLaunchKernelGGL(k_rhs_mean, dim3(1), dim3(WV_RED), 0, st, s, 1);
  hipLaunchKernelGGL(k_rhs_shift, dim3(nb(s.N, 256)), dim3(256), 0, st, s);
}

// ops 6 + 7 with an FP64 dot product: alpha = rho / (p . A p)
__global__ void __launch_bounds__(WV_RED) k_dot_alpha64(Ctx s, int final) {
  if (!solving(s)) return;
  int64_t N = s.N;
  if (!final) {
    int64_t i = GID();
    d4 v = i < N ? mkd((double)s.work[2 * N + i] * (double)s.work[3 * N + i], 0, 0, 0) : mkd(0, 0, 0, 0);
    v = group_reduce_d(v); put_partial(s, v); return;
  }
  double dotp = global_reduce_d(s.partial64, (N + WV_RED - 1) / WV_RED).x;
  if (threadIdx.x == 0) {
    if (!(dotp > 0) || isinf(dotp)) { s.c[7] = 1; return; }
    double alpha = s.dctl[0] / dotp;
    s.c[11] = (float)alpha;
  }
}
void wv_solve_dot_alpha64(const Ctx& s, hipStream_t st) {
  hipLaunchKernelGGL(k_dot_alpha64, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0);
  hipLaunchKernelGGL(k_dot_alpha64, dim3(1), dim3(WV_RED), 0, st, s, 1);
}

// MG op 10 + op 4 / op 9 with FP64 sums (pressure_iterations.glslinc:14-19, 31)
__global__ void __launch_bounds__(WV_RED) k_mg_out64(Ctx s, MGDev m, int field, int phase, int final) {
  if (!mg_enabled(s, phase)) return;
  int64_t N = s.N;
  if (!final) {
    int64_t i = GID(); d4 v = mkd(0, 0, 0, 0);
    if (i < N) {
      float r = s.work[i], z = m.vec[at(m, 0, field, i)], b = s.work[5 * N + i];
      s.work[N + i] = z; if (phase == 0) s.work[2 * N + i] = z;
      v = mkd((double)r * z, (double)r * r, (double)b * b, fabs((double)r));
    }
    v = group_reduce_d(v); put_partial(s, v); return;
  }
  d4 v = global_reduce_d(s.partial64, (N + WV_RED - 1) / WV_RED);
  if (threadIdx.x != 0) return;
  float* c = s.c; double* d = s.dctl;
  if (phase == 0) {
    d[0] = v.x; d[1] = v.z; d[2] = v.y;
    c[9] = (float)v.x; c[10] = (float)v.z; c[13] = (float)sqrt(v.y / (v.z > 1e-30 ? v.z : 1e-30));
    if (converged64(s, v.y, v.z, v.w)) c[6] = 1;
    if (c[58] == 0) c[59] = (float)v.z;
    c[57] = 0;
  } else {
    double beta = v.x / (d[0] > 1e-30 ? d[0] : 1e-30);
    d[0] = v.x; d[2] = v.y;
    c[12] = (float)beta; c[9] = (float)v.x; c[8]++; c[18]++;
    double rel = sqrt(v.y / (d[1] > 1e-30 ? d[1] : 1e-30));
    c[13] = (float)rel;
    if (isnan(rel) || isinf(rel)) c[7] = 2;
    else if (converged64(s, v.y, d[1], v.w)) c[6] = 1;
    else if (c[8] > 4096) c[7] = 3;
  }
}
void wv_launch_mg_out64(const Ctx& s, const MGDev& m, int field, int phase, hipStream_t st) {
  hipLaunchKernelGGL(k_mg_out64, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, m, field, phase, 0);
  hipLaunchKernelGGL(k_mg_out64, dim3(1), dim3(WV_RED), 0, st, s, m, field, phase, 1);
}

// True residual r = b - A p in FP64 from the FP32 pressure and coefficients
// (pressure_core.glslinc:120-125 operator).  dctl[3] = |r|^2, dctl[4] = |b|^2,
// dctl[5] = max|r|.  With `replace`, a converged flag the true residual does not
// confirm is withdrawn, r <- r_true, and the next V-cycle restarts CG (c[5]).
// The pressure is the FP64 accumulator (parts bit 1) plus the FP32 candidate
// pressure (parts bit 0).
__device__ __forceinline__ double pressure64(const Ctx& s, int cb, int64_t i, int parts) {
  return ((parts & 2) ? s.pacc[i] : 0.0) + ((parts & 1) ? (double)getv(s, cb, P_(s) + i) : 0.0);
}
__device__ __forceinline__ double true_residual_at(const Ctx& s, int64_t i, int parts) {
  I3 q = coord(i, dims(s)); int64_t N = s.N;
  if (!wet(s, q)) return 0;
  int cb = candidate(s);
  double here = pressure64(s, cb, i, parts), ap = 0; int face = 0;
  for (int axis = 0; axis < 3; axis++) for (int sign = -1; sign <= 1; sign += 2) {
    I3 nb = q; addc(nb, axis, sign); nb = wrap_cell(s, nb);
    double other = 0;
    if (!outside(s, nb) && wet(s, nb)) other = pressure64(s, cb, index3(nb, dims(s)), parts);
    ap += (double)s.work[(int64_t)(6 + face) * N + i] * (here - other); face++;
  }
  return (double)s.work[5 * N + i] - ap;
}
__global__ void __launch_bounds__(WV_RED) k_true_residual64(Ctx s, int final, int replace, int parts) {
  int64_t N = s.N;
  if (!final) {
    int64_t i = GID(); d4 v = mkd(0, 0, 0, 0);
    if (i < N) { double r = true_residual_at(s, i, parts), b = s.work[5 * N + i]; v = mkd(r * r, b * b, 0, fabs(r)); }
    v = group_reduce_d(v); put_partial(s, v); return;
  }
  d4 g = global_reduce_d(s.partial64, (N + WV_RED - 1) / WV_RED);
  if (threadIdx.x != 0) return;
  s.dctl[3] = g.x; s.dctl[4] = g.y; s.dctl[5] = g.w;
  if (replace && s.c[6] > .5f && s.c[7] < .5f && !converged64(s, g.x, s.dctl[1] > 0 ? s.dctl[1] : g.y, g.w)) {
    s.c[6] = 0; s.dctl[6] = 1;  // restart requested
  } else s.dctl[6] = 0;
}
// restart: pacc += p, p = 0 (every cell, before any residual reads a neighbour)
__global__ void __launch_bounds__(256) k_fold_pressure(Ctx s) {
  if (s.dctl[6] < .5) return;
  int64_t i = GID(); if (i >= s.N) return;
  int cb = candidate(s); int64_t k = P_(s) + i;
  s.pacc[i] += (double)getv(s, cb, k);
  putv(s, cb, k, 0);
}
__global__ void __launch_bounds__(256) k_replace_residual(Ctx s) {
  if (s.dctl[6] < .5) return;
  int64_t i = GID(); if (i >= s.N) return;
  s.work[i] = (float)true_residual_at(s, i, 3);
}
void wv_solve_true_residual64(const Ctx& s, const MGDev& m, bool replace, int parts, hipStream_t st) {
  hipLaunchKernelGGL(k_true_residual64, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0, replace ? 1 : 0, parts);
  hipLaunchKernelGGL(k_true_residual64, dim3(1), dim3(WV_RED), 0, st, s, 1, replace ? 1 : 0, parts);
  if (replace) {
    hipLaunchKernelGGL(k_fold_pressure, dim3(nb(s.N, 256)), dim3(256), 0, st, s);
    hipLaunchKernelGGL(k_replace_residual, dim3(nb(s.N, 256)), dim3(256), 0, st, s);
  }
}

// end of a solve: pacc <- pacc + p (the FP64 solution), p <- its FP32 rounding
__global__ void __launch_bounds__(256) k_solve_finish(Ctx s) {
  int64_t i = GID(); if (i >= s.N) return;
  int cb = candidate(s); int64_t k = P_(s) + i;
  double t = s.pacc[i] + (double)getv(s, cb, k);
  s.pacc[i] = t;
  putv(s, cb, k, (float)t);
}
void wv_solve_finish(const Ctx& s, hipStream_t st) {
  hipLaunchKernelGGL(k_solve_finish, dim3(nb(s.N, 256)), dim3(256), 0, st, s);
}
