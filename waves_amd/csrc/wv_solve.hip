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
#include "wv_netdev.h"

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
#pragma unroll 8
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
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_rhs_mean(Ctx s, int) {
  if (!FINAL) {
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
  hipLaunchKernelGGL(k_rhs_mean<false>, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0);
  hipLaunchKernelGGL(k_rhs_mean<true>, dim3(1), dim3(WV_RED), 0, st, s, 1);
  hipLaunchKernelGGL(k_rhs_shift, dim3(nb(s.N, 256)), dim3(256), 0, st, s);
}

// ops 6 + 7 with an FP64 dot product: alpha = rho / (p . A p)
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_dot_alpha64(Ctx s, int) {
  if (!solving(s)) return;
  int64_t N = s.N;
  if (!FINAL) {
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
  hipLaunchKernelGGL(k_dot_alpha64<false>, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0);
  hipLaunchKernelGGL(k_dot_alpha64<true>, dim3(1), dim3(WV_RED), 0, st, s, 1);
}

// MG op 10 + op 4 / op 9 with FP64 sums (pressure_iterations.glslinc:14-19, 31)
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_mg_out64(Ctx s, MGDev m, int field, int phase, int) {
  if (!mg_enabled(s, phase)) return;
  int64_t N = s.N;
  if (!FINAL) {
    int64_t i = GID(); d4 v = mkd(0, 0, 0, 0);
    if (i < N) {
      float r = s.work[i], z = m.vec[at(m, 0, field, i)], b = s.work[5 * N + i], zold = s.work[N + i];
      s.work[N + i] = z; if (phase == 0) s.work[2 * N + i] = z;
      // phase 1 keeps |b|^2 from phase 0; the flexible form needs r . z_old instead
      double third = (phase == 1 && s.flex) ? (double)r * zold : (double)b * b;
      v = mkd((double)r * z, (double)r * r, third, fabs((double)r));
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
    // Fletcher-Reeves form, or Polak-Ribiere (r.z - r.z_old) / rho for the flexible solve
    double beta = (s.flex ? v.x - v.z : v.x) / (d[0] > 1e-30 ? d[0] : 1e-30);
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
  hipLaunchKernelGGL(k_mg_out64<false>, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, m, field, phase, 0);
  hipLaunchKernelGGL(k_mg_out64<true>, dim3(1), dim3(WV_RED), 0, st, s, m, field, phase, 1);
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
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_true_residual64(Ctx s, int, int replace, int parts) {
  int64_t N = s.N;
  if (!FINAL) {
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
  hipLaunchKernelGGL(k_true_residual64<false>, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0, replace ? 1 : 0, parts);
  hipLaunchKernelGGL(k_true_residual64<true>, dim3(1), dim3(WV_RED), 0, st, s, 1, replace ? 1 : 0, parts);
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

// Given right-hand side: b already in work[5N] (written by the caller), r = b -
// A p from the candidate pressure (the operator of pressure_core.glslinc:120-125),
// z = r / diag, direction z (pressure_iterations.glslinc:4-13 without the
// divergence).
__global__ void __launch_bounds__(256) k_init_given(Ctx s) {
  int64_t i = GID(); if (i >= s.N) return;
  int64_t N = s.N; I3 q = coord(i, dims(s));
  float r = 0, z = 0, diag = 1;
  if (wet(s, q)) { diag = s.work[12 * N + i]; r = s.work[5 * N + i] - operator_at(s, q, 0); z = r / diag; }
  s.work[i] = r; s.work[N + i] = z; s.work[2 * N + i] = z; s.work[4 * N + i] = diag;
}
void wv_solve_init_given(const Ctx& s, hipStream_t st) {
  hipLaunchKernelGGL(k_init_given, dim3(nb(s.N, 256)), dim3(256), 0, st, s);
}

// ------------------------------------------------- learned warm start ----
// Network input per cell: F, r / rms_wet(r), p / rms_wet(p), wet, 0 x 4 (bf16).
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_net_scale(Ctx s, int) {
  int64_t N = s.N;
  if (!FINAL) {
    int64_t i = GID(); d4 v = mkd(0, 0, 0, 0);
    if (i < N && wet(s, coord(i, dims(s)))) {
      double r = s.work[i], p = getv(s, candidate(s), P_(s) + i);
      v = mkd(r * r, p * p, 1, 0);
    }
    v = group_reduce_d(v); put_partial(s, v); return;
  }
  d4 g = global_reduce_d(s.partial64, (N + WV_RED - 1) / WV_RED);
  if (threadIdx.x == 0) {
    double n = g.z > 0 ? g.z : 1;
    s.dctl[9] = sqrt(g.x / n); s.dctl[10] = sqrt(g.y / n);
  }
}
__global__ void __launch_bounds__(256) k_net_features(Ctx s, uint16_t* x) {
  int64_t i = GID(); if (i >= s.N) return;
  int cb = candidate(s);
  bool w = wet(s, coord(i, dims(s)));
  double sr = s.dctl[9], sp = s.dctl[10];
  float f0 = getv(s, cb, i);
  float f1 = (w && sr > 0) ? (float)((double)s.work[i] / sr) : 0.f;
  float f2 = (w && sp > 0) ? (float)((double)getv(s, cb, P_(s) + i) / sp) : 0.f;
  uint16_t* o = x + i * 8;
  o[0] = f32_to_bf16(f0); o[1] = f32_to_bf16(f1); o[2] = f32_to_bf16(f2); o[3] = f32_to_bf16(w ? 1.f : 0.f);
  o[4] = 0; o[5] = 0; o[6] = 0; o[7] = 0;
}
// basis vectors: v1 = fine head, v2 = coarse head prolonged by 4 per axis, both
// zero outside the wet cells
__device__ __forceinline__ float ws_v(const Ctx& s, const float* qf, const float* qc, int which, I3 q) {
  q = wrap_cell(s, q);
  if (!wet(s, q)) return 0.f;
  I3 d = dims(s);
  if (which == 0) return qf[index3(q, d)];
  int n2x = ((d.x + 1) / 2 + 1) / 2, n2y = ((d.y + 1) / 2 + 1) / 2;
  return qc[(q.x >> 2) + (int64_t)n2x * ((q.y >> 2) + (int64_t)n2y * (q.z >> 2))];
}
__device__ __forceinline__ float ws_apply(const Ctx& s, const float* qf, const float* qc, int which, I3 q, int64_t i) {
  float here = ws_v(s, qf, qc, which, q), result = 0; int face = 0;
  for (int axis = 0; axis < 3; axis++) for (int sign = -1; sign <= 1; sign += 2) {
    I3 nb = q; addc(nb, axis, sign);
    result += s.work[(int64_t)(6 + face) * s.N + i] * (here - ws_v(s, qf, qc, which, nb)); face++;
  }
  return result;
}
// Galerkin correction on span{v1, v2}: minimise the A-norm error of p + a v1 + c v2,
// [v1 A v1, v1 A v2; v2 A v1, v2 A v2] [a; c] = [v1 . r; v2 . r], in FP64
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_ws_dots(Ctx s, const float* qf, const float* qc, int) {
  int64_t N = s.N, groups = (N + WV_RED - 1) / WV_RED;
  if (!FINAL) {
    int64_t i = GID(); d4 v = mkd(0, 0, 0, 0), u = mkd(0, 0, 0, 0);
    if (i < N) {
      I3 q = coord(i, dims(s));
      if (wet(s, q)) {
        double v1 = ws_v(s, qf, qc, 0, q), v2 = ws_v(s, qf, qc, 1, q), r = s.work[i];
        double a1 = ws_apply(s, qf, qc, 0, q, i), a2 = ws_apply(s, qf, qc, 1, q, i);
        v = mkd(v1 * a1, v1 * a2, v2 * a2, 0); u = mkd(v1 * r, v2 * r, 0, 0);
      }
    }
    v = group_reduce_d(v); u = group_reduce_d(u);
    if (threadIdx.x == 0) { s.partial64[blockIdx.x] = make_double4(v.x, v.y, v.z, v.w); s.partial64[groups + 1 + blockIdx.x] = make_double4(u.x, u.y, u.z, u.w); }
    return;
  }
  d4 a = global_reduce_d(s.partial64, groups);
  d4 b = global_reduce_d(s.partial64 + groups + 1, groups);
  if (threadIdx.x != 0) return;
  double A11 = a.x, A12 = a.y, A22 = a.z, b1 = b.x, b2 = b.y, al = 0, ce = 0;
  double det = A11 * A22 - A12 * A12;
  if (A11 > 0 && A22 > 0 && det > 1e-10 * A11 * A22) { al = (b1 * A22 - b2 * A12) / det; ce = (A11 * b2 - A12 * b1) / det; }
  else if (A11 > 0) al = b1 / A11;
  else if (A22 > 0) ce = b2 / A22;
  if (!isfinite(al) || !isfinite(ce)) { al = 0; ce = 0; }
  s.dctl[11] = al; s.dctl[12] = ce;
}
__global__ void __launch_bounds__(256) k_ws_update(Ctx s, const float* qf, const float* qc) {
  int64_t i = GID(); if (i >= s.N) return;
  I3 q = coord(i, dims(s)); if (!wet(s, q)) return;
  double v = s.dctl[11] * ws_v(s, qf, qc, 0, q) + s.dctl[12] * ws_v(s, qf, qc, 1, q);
  int cb = candidate(s); int64_t k = P_(s) + i;
  putv(s, cb, k, (float)((double)getv(s, cb, k) + v));
}
void wv_solve_learned_start(const Ctx& s, wv_net* net, hipStream_t st) {
  hipLaunchKernelGGL(k_net_scale<false>, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0);
  hipLaunchKernelGGL(k_net_scale<true>, dim3(1), dim3(WV_RED), 0, st, s, 1);
  hipLaunchKernelGGL(k_net_features, dim3(nb(s.N, 256)), dim3(256), 0, st, s, wv_net_input_dev(net));
  wv_net_forward_dev(net, st);
  const float* qf = wv_net_qf_dev(net); const float* qc = wv_net_qc_dev(net);
  hipLaunchKernelGGL(k_ws_dots<false>, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, qf, qc, 0);
  hipLaunchKernelGGL(k_ws_dots<true>, dim3(1), dim3(WV_RED), 0, st, s, qf, qc, 1);
  hipLaunchKernelGGL(k_ws_update, dim3(nb(s.N, 256)), dim3(256), 0, st, s, qf, qc);
  hipLaunchKernelGGL(k_init_given, dim3(nb(s.N, 256)), dim3(256), 0, st, s);   // r = b - A p
}

// Global liquid volume sum_i F_i h_x h_y h_z of one bank in FP64 (the FP32
// volume of the acceptance op, pressure_pass.glslinc:78-103, is c[17]);
// deterministic: fixed 128-lane groups, then one group -> dctl[13].
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_volume64(Ctx s, int bank, int) {
  int64_t N = s.N;
  if (!FINAL) {
    int64_t i = GID();
    d4 v = mkd(i < N ? (double)getv(s, bank, i) : 0.0, 0, 0, 0);
    v = group_reduce_d(v); put_partial(s, v); return;
  }
  d4 g = global_reduce_d(s.partial64, (N + WV_RED - 1) / WV_RED);
  if (threadIdx.x == 0) s.dctl[13] = g.x * ((double)s.h[0] * (double)s.h[1] * (double)s.h[2]);
}
void wv_solve_volume64(const Ctx& s, int bank, hipStream_t st) {
  hipLaunchKernelGGL(k_volume64<false>, dim3(nb(s.N, WV_RED)), dim3(WV_RED), 0, st, s, bank, 0);
  hipLaunchKernelGGL(k_volume64<true>, dim3(1), dim3(WV_RED), 0, st, s, bank, 1);
}
