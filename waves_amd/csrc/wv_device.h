// Device-side state and per-cell/per-face helpers of the gfx950 solver.
//
// Data layout in HBM is the reference's staggered MAC layout
// (water/shaders/compute/pressure_core.glslinc:2-11, 28-36, coupled_motion.glslinc:8-14):
//   bank A/B : [F N][U (nx+1)ny nz][V nx(ny+1)nz][W nx ny(nz+1)][P N], FP32, x fastest
//   geo      : float2 (signed distance phi, kind -1 solid / 0 air / 1 wet) per cell
//   work     : 14 cell fields (r, z, p-dir, Ap, diag, b, 6 face coefficients, diag, saved p)
//   motion   : [scratch F N][c_i N][flux F][forward F][source 2F][valid 2F][pmask 2N][trace cache 6F]
//   control  : c[96], the device-resident trial state machine (acceptance, PCG, ledgers)
// Every helper below restates the per-invocation arithmetic of the reference in
// the same operation order, and the library is compiled with -ffp-contract=off,
// so rational results are bit-identical to the reference's FP32 semantics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define WV_NCTRL 96
#define WV_RED 128  // lanes of one reduction group (the reference's workgroup)

struct Ctx {
  // geometry and parameters (the reference's push constants, pressure_core.glslinc:34-43)
  int nx, ny, nz;
  int64_t N, NU, NV, NW, NF, stride;
  float h[3], dt_max, o[3], eta, perx, slope, corr;
  float shape, gmult, max_face_cfl, hist_cap;
  float wave[4];
  float tol, linear_air, stokes2;
  int flags;
  float forc[4];
  float right_abs, pert_amp;
  uint32_t seed;
  // device buffers
  float* bank[2];
  float* work;
  float2* geo;
  float* c;
  float4* plane;
  float* motion;
  float4* history;
  float4* chist;
  float4* probe;
  float4* img_planes;
  float* img_kinds;
  float4* partial;
  double4* partial64;  // FP64 per-group partials (mixed-precision solve)
  double* dctl;        // FP64 scalars of the mixed-precision solve: rho, b2, r2, true r2, mean b
  double* pacc;        // FP64 pressure accumulator of the mixed-precision solve (N)
  float* aux;          // 2 N scratch of the learned warm start (A v1, A v2)
  int flex;            // flexible (Polak-Ribiere) CG coefficient in the FP64 V-cycle output
  float4* vpart;       // 4 G + 4, G = ceil(N / 128): the VOF update's group partials per sweep and the
                       // sweep headers whose ledger sums the acceptance finaliser adds (batch program)
  int* mq;             // faces whose MacCormack trace needs the general sampling path (NF)
  unsigned* mq_count;  // queue lengths: [0] forward trace (op 27), [1] correction (op 28), [2] exact
                       // distance; then 8 sub-queue counters per queue slot (4 slots)
  signed char* reflect;  // mirror code of the wall search per (component, y, x), 3 (ny + 1) (nx + 1)
  const float* quota;  // device copy of the request quota (steps >= 0, -deadline < 0)
  // pressure-section timing on the device clock (telemetry, no reference counterpart)
  unsigned long long* tstate;      // [0] section start, [1] section active, [2] trial ns, [3] ns since last acceptance,
                                   // [4] accepted steps seen, [5] trial count seen, [6] sections of the trial
  unsigned long long* step_times;  // per accepted step (history capacity rows): trial ns, all-trials ns, trials, sections
  float tick_ns;                   // device wall-clock period [ns]
};

struct I3 { int x, y, z; };
__device__ __forceinline__ I3 mk(int x, int y, int z) { I3 r; r.x = x; r.y = y; r.z = z; return r; }
// Component access by a run-time axis through selects: a reference to a
// struct member would take its address and push the struct to scratch memory.
__device__ __forceinline__ void addc(I3& q, int a, int v) { q.x += a == 0 ? v : 0; q.y += a == 1 ? v : 0; q.z += a == 2 ? v : 0; }
__device__ __forceinline__ void setc(I3& q, int a, int v) { q.x = a == 0 ? v : q.x; q.y = a == 1 ? v : q.y; q.z = a == 2 ? v : q.z; }
__device__ __forceinline__ int compc(const I3& q, int a) { return a == 0 ? q.x : (a == 1 ? q.y : q.z); }
struct F3 { float x, y, z; };
__device__ __forceinline__ F3 mkf(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ void addc(F3& q, int a, float v) { if (a == 0) q.x += v; else if (a == 1) q.y += v; else q.z += v; }
__device__ __forceinline__ void setc(F3& q, int a, float v) { q.x = a == 0 ? v : q.x; q.y = a == 1 ? v : q.y; q.z = a == 2 ? v : q.z; }
__device__ __forceinline__ float compc(const F3& q, int a) { return a == 0 ? q.x : (a == 1 ? q.y : q.z); }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(F3 a) { return sqrtf(dot3(a, a)); }
// GLSL min/max semantics (second operand wins only if strictly smaller/larger)
__device__ __forceinline__ float gmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float gmax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
__device__ __forceinline__ int iclamp(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
__device__ __forceinline__ float fract(float x) { return x - floorf(x); }
__device__ __forceinline__ float gsign(float x) { return x > 0 ? 1.0f : (x < 0 ? -1.0f : 0.0f); }
__device__ __forceinline__ int pidx(int q, int n) {
  if ((unsigned)q < (unsigned)n) return q;   // in range: the modulo is the identity
  return q >= 0 ? q % n : n - 1 - ((-1 - q) % n);
}

__device__ __forceinline__ float hc(const Ctx& s, int a) { return s.h[a]; }
__device__ __forceinline__ int64_t U_(const Ctx& s) { return s.N; }
__device__ __forceinline__ int64_t V_(const Ctx& s) { return s.N + s.NU; }
__device__ __forceinline__ int64_t W_(const Ctx& s) { return s.N + s.NU + s.NV; }
__device__ __forceinline__ int64_t P_(const Ctx& s) { return s.N + s.NU + s.NV + s.NW; }
__device__ __forceinline__ int64_t index3(I3 q, I3 n) { return (int64_t)q.x + (int64_t)n.x * (q.y + (int64_t)n.y * q.z); }
__device__ __forceinline__ I3 dims(const Ctx& s) { return mk(s.nx, s.ny, s.nz); }
__device__ __forceinline__ I3 coord(int64_t i, I3 n) {
  // 32-bit unsigned division (about a fifth of the 64-bit sequence) whenever
  // the index and a plane of the grid fit, which is every grid here
  if ((uint64_t)i < 0x100000000ull && (uint64_t)n.x * (uint64_t)n.y < 0x100000000ull) {
    uint32_t u = (uint32_t)i, xy = (uint32_t)n.x * (uint32_t)n.y;
    uint32_t z = u / xy, r = u - z * xy, y = r / (uint32_t)n.x;
    return mk((int)(r - y * (uint32_t)n.x), (int)y, (int)z);
  }
  int64_t xy = (int64_t)n.x * n.y;
  int z = (int)(i / xy); int64_t r = i - (int64_t)z * xy; int y = (int)(r / n.x);
  return mk((int)(r - (int64_t)y * n.x), y, z);
}
__device__ __forceinline__ float getv(const Ctx& s, int bank, int64_t i) { return s.bank[bank][i]; }
__device__ __forceinline__ void putv(const Ctx& s, int bank, int64_t i, float v) { s.bank[bank][i] = v; }
__device__ __forceinline__ float timestep(const Ctx& s) { return s.c[20]; }
__device__ __forceinline__ int accepted(const Ctx& s) { return (int)s.c[0]; }
__device__ __forceinline__ int candidate(const Ctx& s) { return 1 - accepted(s); }
// pressure_core.glslinc:66-72; `export_op` is true only for the accepted export (op 19)
__device__ __forceinline__ int geometry_bank(const Ctx& s, bool export_op) { return (export_op || s.c[4] < .5f) ? accepted(s) : candidate(s); }

// pressure_core.glslinc:73-83
__device__ __forceinline__ float eta(const Ctx& s, I3 q) {
  if (s.shape == 5) {
    float phase = s.wave[1] * (s.o[0] + ((float)q.x + .5f) * s.h[0]);
    float sigma = tanhf(s.wave[1] * s.wave[3]);
    float second = s.stokes2 > .5f ? s.wave[1] * s.wave[0] * s.wave[0] * (3 - sigma * sigma) / (4 * sigma * sigma * sigma) : 0;
    return s.eta + s.wave[0] * cosf(phase) + second * cosf(2 * phase);
  }
  return s.eta + s.corr * cosf(6.28318530718f * ((float)q.x + 0.5f) / (float)s.nx);
}
__device__ __forceinline__ F3 center(const Ctx& s, I3 q) {
  return mkf(s.o[0] + ((float)q.x + 0.5f) * s.h[0], s.o[1] + ((float)q.y + 0.5f) * s.h[1], s.o[2] + ((float)q.z + 0.5f) * s.h[2]);
}
__device__ __forceinline__ I3 wrap_cell(const Ctx& s, I3 q) {
  q.z = pidx(q.z, s.nz); if (s.perx > 0.5f) q.x = pidx(q.x, s.nx); return q;
}
__device__ __forceinline__ bool outside(const Ctx& s, I3 q) { return q.x < 0 || q.x >= s.nx || q.y < 0 || q.y >= s.ny; }
// pressure_core.glslinc:87-96; x < 0 is the open ocean of the coastal program
__device__ __forceinline__ bool solid_formula(const Ctx& s, I3 q) {
  if (s.slope > 0.0f) {
    F3 cc = center(s, q);
    float bed = -0.45f + gmax(0.0f, cc.x - 1.5f) / s.slope;
    return cc.y <= bed;
  }
  return false;
}
__device__ __forceinline__ bool solid(const Ctx& s, I3 q) {
  q = wrap_cell(s, q);
  if (q.y < 0 || q.x >= s.nx) return true;
  if (q.y >= s.ny) return false;
  return solid_formula(s, q);
}
__device__ __forceinline__ bool wet(const Ctx& s, I3 q) { q = wrap_cell(s, q); return !outside(s, q) && s.geo[index3(q, dims(s))].y > 0.5f; }
__device__ __forceinline__ float phi(const Ctx& s, I3 q) { q = wrap_cell(s, q); return outside(s, q) ? center(s, q).y - eta(s, q) : s.geo[index3(q, dims(s))].x; }
// wet(s, q) ? pressure : 0, the flag and the pressure loaded together
__device__ __forceinline__ float press(const Ctx& s, I3 q) {
  q = wrap_cell(s, q);
  const bool out = outside(s, q);
  const int64_t i = out ? 0 : index3(q, dims(s));
  const float flag = s.geo[i].y, v = getv(s, candidate(s), P_(s) + i);
  return !out && flag > 0.5f ? v : 0.0f;
}
// The wet flag and the value are loaded together (index 0 for a cell outside
// the grid) and the flag selects: wet(s, q) ? value : 0.
__device__ __forceinline__ float vector_at(const Ctx& s, I3 q, int which) {
  q = wrap_cell(s, q);
  const bool out = outside(s, q);
  const int64_t i = out ? 0 : index3(q, dims(s));
  const float flag = s.geo[i].y;
  const float v = which == 0 ? getv(s, candidate(s), P_(s) + i) : s.work[2 * s.N + i];
  return !out && flag > 0.5f ? v : 0.0f;
}
// pressure_core.glslinc:120-125: masked 7-point operator
__device__ __forceinline__ float operator_at(const Ctx& s, I3 q, int which) {
  float result = 0, here = vector_at(s, q, which); int64_t i = index3(q, dims(s));
  int face = 0;
  for (int axis = 0; axis < 3; axis++) for (int sign = -1; sign <= 1; sign += 2) {
    I3 nb = q; addc(nb, axis, sign); float k = s.work[(int64_t)(6 + face) * s.N + i];
    result += k * (here - vector_at(s, nb, which)); face++;
  }
  return result;
}
// pressure_core.glslinc:100-109: ghost-fluid face distance
__device__ __forceinline__ float distance_face(const Ctx& s, I3 l, I3 r, int axis) {
  float h = s.h[axis];
  if (axis == 1 && (l.y >= s.ny || r.y >= s.ny)) return .5f * h;
  bool wl = wet(s, l), wr = wet(s, r);
  if (wl && wr) return h;
  float pl = fabsf(phi(s, l)), pr = fabsf(phi(s, r));
  return h * gmax((wl ? pl : pr) / gmax(pl + pr, 1e-30f), 1e-6f);
}
// pressure_core.glslinc:110-118
__device__ __forceinline__ float coefficient(const Ctx& s, I3 q, I3 nb, int axis) {
  if (solid(s, nb)) return 0.0f;
  if (axis == 0 && nb.x < 0) return 0.0f;
  return 1.0f / (s.h[axis] * distance_face(s, q, nb, axis));
}
// pressure_core.glslinc:129-133: periodic duplicates address the canonical face
__device__ __forceinline__ int64_t face_index(const Ctx& s, I3 q, int axis) {
  I3 n = dims(s);
  if (axis == 2 && q.z == n.z) q.z = 0;
  if (axis == 0 && s.perx > .5f && q.x == n.x) q.x = 0;
  addc(n, axis, 1);
  return (axis == 0 ? U_(s) : (axis == 1 ? V_(s) : W_(s))) + index3(q, n);
}
__device__ __forceinline__ float divergence(const Ctx& s, I3 q) {
  int cb = candidate(s); float d = 0;
  for (int axis = 0; axis < 3; axis++) {
    I3 r = q; addc(r, axis, 1);
    d += (getv(s, cb, face_index(s, r, axis)) - getv(s, cb, face_index(s, q, axis))) / s.h[axis];
  }
  return d;
}
// pressure_core.glslinc:156-179
__device__ __forceinline__ bool pressure_setup(const Ctx& s) { return s.c[5] > .5f || s.c[57] == 2; }
__device__ __forceinline__ float pressure_scaled_norm(const Ctx& s, float sq) { float dt = timestep(s); return dt * dt / 1000 * sqrtf(gmax(sq, 0)); }
__device__ __forceinline__ bool pressure_norm_converged(const Ctx& s, float r2, float b2) {
  return r2 <= s.tol * s.tol * b2 || (pressure_scaled_norm(s, b2) <= 1e-5f && pressure_scaled_norm(s, r2) <= 1e-8f);
}
__device__ __forceinline__ bool residual_converged(const Ctx& s, float4 sums, float b2) {
  float dt = timestep(s);
  return pressure_norm_converged(s, sums.y, b2) && dt * dt / 1000 * sums.w < 5e-5f;
}
__device__ __forceinline__ bool solving(const Ctx& s) { return s.c[4] > 0.5f && s.c[6] < 0.5f && s.c[7] < 0.5f; }

// motion layout, coupled_motion.glslinc:8-14, 151
__device__ __forceinline__ int64_t flux_base(const Ctx& s) { return 2 * s.N; }
__device__ __forceinline__ int64_t forward_base(const Ctx& s) { return 2 * s.N + s.NF; }
__device__ __forceinline__ int64_t source_base(const Ctx& s, int b) { return 2 * s.N + 2 * s.NF + b * s.NF; }
__device__ __forceinline__ int64_t valid_base(const Ctx& s, int b) { return 2 * s.N + 4 * s.NF + b * s.NF; }
__device__ __forceinline__ int64_t pmask_base(const Ctx& s, int b) { return 2 * s.N + 6 * s.NF + b * s.N; }
__device__ __forceinline__ int64_t trace_base(const Ctx& s) { return 4 * s.N + 6 * s.NF; }
__device__ __forceinline__ int64_t face_local(const Ctx& s, I3 q, int axis) { return face_index(s, q, axis) - U_(s); }
__device__ __forceinline__ void unpack_face(const Ctx& s, int64_t f, I3& q, int& axis) {
  axis = f < s.NU ? 0 : (f < s.NU + s.NV ? 1 : 2);
  int64_t base = axis == 0 ? 0 : (axis == 1 ? s.NU : s.NU + s.NV);
  I3 n = dims(s); addc(n, axis, 1); q = coord(f - base, n);
}

// ---- wavefront-shuffle reductions -------------------------------------------
// One reduction group is 128 lanes = two 64-wide wavefronts.  The combine tree
// pairs lane l with l+64 (through LDS), then l with l+32 ... l+1 (DPP/shuffle
// inside the wavefront): exactly the pairing of the reference's 128-lane LDS
// tree (pressure_core.glslinc:135), so float sums come out bit-identical.
enum RedMode { RED_SUM = 0, RED_PRESSURE = 1, RED_BOUNDS = 2, RED_DIAG = 3 };
template <int MODE>
__device__ __forceinline__ float4 red_combine(float4 a, float4 b) {
  float4 r;
  if (MODE == RED_SUM) { r.x = a.x + b.x; r.y = a.y + b.y; r.z = a.z + b.z; r.w = a.w + b.w; }
  else if (MODE == RED_PRESSURE) { r.x = a.x + b.x; r.y = a.y + b.y; r.z = a.z + b.z; r.w = gmax(a.w, b.w); }
  else if (MODE == RED_BOUNDS) { r.x = gmin(a.x, b.x); r.y = gmax(a.y, b.y); r.z = a.z + b.z; r.w = a.w + b.w; }
  else { r.x = gmax(a.x, b.x); r.y = gmax(a.y, b.y); r.z = gmax(a.z, b.z); r.w = a.w + b.w; }
  return r;
}
__device__ __forceinline__ float4 shfl_down4(float4 v, int off) {
  v.x = __shfl_down(v.x, off, 64); v.y = __shfl_down(v.y, off, 64);
  v.z = __shfl_down(v.z, off, 64); v.w = __shfl_down(v.w, off, 64);
  return v;
}
// Reduce one 128-lane group (blockDim.x must be 128); the result is valid in lane 0.
template <int MODE>
__device__ __forceinline__ float4 group_reduce(float4 v) {
  __shared__ float4 upper[64];
  int lid = threadIdx.x;
  if (lid >= 64) upper[lid - 64] = v;
  __syncthreads();
  if (lid < 64) {
    v = red_combine<MODE>(v, upper[lid]);
    for (int off = 32; off > 0; off >>= 1) v = red_combine<MODE>(v, shfl_down4(v, off));
  }
  __syncthreads();
  return v;
}
// pressure_core.glslinc:146-155: 128 lanes, strided accumulation then the tree.
// MAXW: the w component is a max (pressure partials) instead of a sum.
template <bool MAXW>
__device__ __forceinline__ float4 global_reduce(const float4* partial, int64_t groups) {
  int lid = threadIdx.x; float4 v = make_float4(0, 0, 0, 0);
  // unrolled: the loads run ahead, each lane still adds its partials in index order
#pragma unroll 8
  for (int64_t k = lid; k < groups; k += WV_RED) {
    float4 p = partial[k];
    v.x += p.x; v.y += p.y; v.z += p.z; v.w = MAXW ? gmax(v.w, p.w) : v.w + p.w;
  }
  return group_reduce<MAXW ? RED_PRESSURE : RED_SUM>(v);
}
