// Per-step kernels of the gfx950 solver: initialisation, trial control, split
// geometric VOF, limited MacCormack, coastal relaxation, geometry, PCG vector
// ops, projection, extension and acceptance.  Each kernel restates one op (or
// one op and its single-group finaliser, as a second one-group launch) of
// water/shaders/compute/{pressure_pass,pressure_iterations,coupled_motion,
// coastal_pass}.glslinc.  Reductions run in 128-lane groups with wavefront
// shuffles so that their float sums match the reference's workgroup trees.
#include "wv_motion.h"
#include "wv_kernels.h"
#include "wv_mgdev.h"

#define GID() ((int64_t)blockIdx.x * blockDim.x + threadIdx.x)

// ---------------------------------------------------------------- init ops --
// pressure_pass.glslinc:3-19 (op 0)
__global__ void __launch_bounds__(256) k_op0(Ctx s) {
  int64_t i = GID();
  if (i < s.stride) { s.bank[0][i] = 0; s.bank[1][i] = 0; }
  if (i < WV_NCTRL) s.c[i] = 0;
  if (i < s.N) {
    I3 q = coord(i, dims(s)); float ph = center(s, q).y - eta(s, q);
    bool wall = solid(s, q); s.geo[i] = make_float2(ph, wall ? -1.0f : (ph < 0.0f ? 1.0f : 0.0f));
    float f = wall ? 0.0f : gclamp(0.5f - ph / s.h[1], 0.0f, 1.0f); s.bank[0][i] = f; s.bank[1][i] = f;
  }
}
// coupled_motion.glslinc:234-270 (op 40): fixtures, dt, gravity multiplier
__global__ void __launch_bounds__(256) k_op40(Ctx s) {
  int64_t i = GID(); int acc = accepted(s);
  if (i < s.N) {
    I3 q = coord(i, dims(s));
    float lo[3] = {(float)q.x * s.h[0], (float)q.y * s.h[1], (float)q.z * s.h[2]};
    float hi[3] = {lo[0] + s.h[0], lo[1] + s.h[1], lo[2] + s.h[2]};
    float f = getv(s, acc, i);
    if (s.shape > .5f && s.shape < 2.5f) {
      float lower[3], upper[3];
      if (s.shape < 1.5f) { lower[0] = 0; lower[1] = 0; lower[2] = 0; upper[0] = .28f; upper[1] = .3f; upper[2] = 100; }
      else { lower[0] = .22f; lower[1] = .28f; lower[2] = .06f; upper[0] = .42f; upper[1] = .48f; upper[2] = .22f; }
      f = 1; for (int k = 0; k < 3; k++) f *= gclamp((gmin(hi[k], upper[k]) - gmax(lo[k], lower[k])) / s.h[k], 0, 1);
      if (s.shape > 1.5f) f += gclamp((.08f - lo[1]) / s.h[1], 0, 1);
    }
    if (s.shape == 4) {
      f = 0;
      for (int block = 0; block < 2; block++) {
        float lower[3] = {block == 0 ? .08f : .36f, .12f, 0}, upper[3] = {block == 0 ? .24f : .52f, .32f, 100};
        float o = 1; for (int k = 0; k < 3; k++) o *= gclamp((gmin(hi[k], upper[k]) - gmax(lo[k], lower[k])) / s.h[k], 0, 1);
        f += o;
      }
    }
    s.bank[0][i] = f; s.bank[1][i] = f;
  }
  if (i < s.NF) {
    I3 q; int axis; unpack_face(s, i, q, axis);
    float velocity = s.shape == 3 && axis == 0 && !wall_face(s, q, axis) && center(s, q).y < s.eta + 3 * s.h[1] ? .25f : 0;
    if (s.shape == 4 && axis == 0 && !wall_face(s, q, axis)) velocity = (float)q.x * s.h[0] < .3f ? 2 : -2;
    if (s.shape == 5) {
      F3 x = mkf(((float)q.x + .5f) * s.h[0] + s.o[0], ((float)q.y + .5f) * s.h[1] + s.o[1], ((float)q.z + .5f) * s.h[2] + s.o[2]);
      addc(x, axis, -(.5f * s.h[axis]));
      if (x.y < eta(s, q) + 3 * s.h[1] && !wall_face(s, q, axis)) {
        if (axis == 0) { F3 p = x, m = x; p.y += .5f * s.h[1]; m.y -= .5f * s.h[1]; velocity = (linear_streamfunction(s, p) - linear_streamfunction(s, m)) / s.h[1]; }
        if (axis == 1) { F3 p = x, m = x; p.x += .5f * s.h[0]; m.x -= .5f * s.h[0]; velocity = -(linear_streamfunction(s, p) - linear_streamfunction(s, m)) / s.h[0]; }
      }
    }
    s.bank[0][U_(s) + i] = velocity; s.bank[1][U_(s) + i] = velocity;
  }
  if (i == 0) { s.c[20] = s.dt_max; s.c[31] = s.gmult; }
}
// Seeded alongshore perturbation of the still surface (benchmark input; no
// reference op): F = clamp(1/2 - (y - eta - a*u(z))/dy), u(z) in [-1, 1).
__device__ __forceinline__ uint32_t wv_hash32(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16; return x;
}
__global__ void __launch_bounds__(256) k_perturb(Ctx s) {
  int64_t i = GID(); if (i >= s.N || s.pert_amp == 0) return;
  I3 q = coord(i, dims(s)); if (solid(s, q)) return;
  uint32_t hsh = wv_hash32(s.seed * 0x9e3779b9U + (uint32_t)q.z * 0x85ebca6bU + 0x632be5abU);
  float u = (float)(hsh >> 8) * (1.0f / 8388608.0f) - 1.0f;
  float ph = center(s, q).y - (eta(s, q) + s.pert_amp * u);
  float f = gclamp(0.5f - ph / s.h[1], 0.0f, 1.0f); s.bank[0][i] = f; s.bank[1][i] = f;
}

// ------------------------------------------------------- geometry (17,18,16) --
// Work queues (Ctx::mq): a first pass handles the cells or faces on a cheap
// path and appends the others; a second pass works the queue in dense waves.
// Slots: 0/1 MacCormack ops 27/28, 2 distance.
__global__ void k_queue_reset(Ctx s, int slot) {
  if (blockIdx.x == 0 && threadIdx.x < 9) { if (threadIdx.x == 8) s.mq_count[slot] = 0; else s.mq_count[4 + 8 * slot + threadIdx.x] = 0; }
}
// Block-aggregated work queues: a launch's items go to eight sub-queues, one
// per XCD of the block order (block b to sub-queue b mod 8, as MI355X hands
// workgroups to its eight XCDs), each with its own counter, and each block
// takes its slots with one atomic.  Thousands of single-address atomics from
// the whole chip no longer serialise the pass.  Sub-queue k holds the items of
// blocks b = k mod 8 at mq[k * cap ...] with cap = ceil(blocks / 8) * block
// size, which bounds its length (mq has room for 8 cap items: it is sized for
// the face count plus 8 blocks).  Every thread of the block makes the call.
__device__ __forceinline__ void queue_append_block(const Ctx& s, bool defer, int64_t i, int slot, int cap) {
  __shared__ unsigned wcount[16], wbase[16], bbase;
  const int lane = __lane_id(), wave = (int)(threadIdx.x >> 6), nw = (int)((blockDim.x + 63) >> 6);
  const unsigned long long ballot = __ballot(defer);
  if (lane == 0) wcount[wave] = (unsigned)__popcll(ballot);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tot = 0;
    for (int w = 0; w < nw; w++) { wbase[w] = tot; tot += wcount[w]; }
    bbase = tot ? atomicAdd(&s.mq_count[4 + 8 * slot + (blockIdx.x & 7)], tot) : 0u;
  }
  __syncthreads();
  if (defer) s.mq[(int64_t)(blockIdx.x & 7) * cap + bbase + wbase[wave] + (unsigned)__popcll(ballot & ((1ull << lane) - 1))] = (int)i;
}
// the sub-queues' lengths (cnt) and their sum; item v of the concatenation
struct QueueView {
  unsigned cnt[8]; int64_t total; int cap;
  __device__ __forceinline__ int item(const Ctx& s, int64_t v) const {
    int k = 0; int64_t base = 0;
    while (k < 7 && v >= base + cnt[k]) { base += cnt[k]; k++; }
    return s.mq[(int64_t)k * cap + (v - base)];
  }
};
__device__ __forceinline__ QueueView queue_view(const Ctx& s, int slot, int cap) {
  QueueView q; q.total = 0; q.cap = cap;
#pragma unroll
  for (int k = 0; k < 8; k++) { q.cnt[k] = s.mq_count[4 + 8 * slot + k]; q.total += q.cnt[k]; }
  if (blockIdx.x == 0 && threadIdx.x == 0) s.mq_count[slot] = (unsigned)q.total;   // telemetry
  return q;
}
constexpr int QUEUE_BLOCKS = 4096;
static inline int64_t red_groups_host(const Ctx& s) { return (s.N + WV_RED - 1) / WV_RED; }
// sub-queue capacity of a launch of `items` threads in blocks of `block`
static inline int queue_cap(int64_t items, int block) { return (int)(((items + block - 1) / block + 7) / 8) * block; }
// MODE 0: op 17 (pressure geometry, gated), 1: op 23 (transport plane), 2: op 19 (export)
template <int MODE>
__device__ __forceinline__ bool reconstruct_gate(const Ctx& s) {
  if (MODE == 0) return !(s.c[4] > .5f && s.c[5] < .5f);
  if (MODE == 1) return s.c[5] > .5f && s.c[7] < .5f;
  return !(s.c[4] > .5f);
}
template <int MODE>
__device__ __forceinline__ void reconstruct_store(const Ctx& s, int64_t i, float4 e) {
  if (MODE == 2) s.img_planes[i] = e; else s.plane[i] = e;
}
// pressure_pass.glslinc:116-130 (ops 17, 19) and coupled_motion.glslinc:50-56 (op 23):
// the PLIC plane of every cell (interface cells: Youngs gradient + bisection)
template <int MODE>
__global__ void __launch_bounds__(256) k_reconstruct(Ctx s) {
  if (!reconstruct_gate<MODE>(s)) return;
  int64_t i = GID(); if (i >= s.N) return;
  I3 q = coord(i, dims(s));
  const int gbank = geometry_bank(s, MODE == 2), source = MODE == 1 ? 1 : 0;
  float f = MODE == 1 ? transport_f(s, q) : geometry_fraction(s, q, gbank);
  float4 e = (f <= 0 || f >= 1) ? make_float4(0, 1, 0, f) : plane_from_gradient(s, q, interface_gradient(s, q, source, gbank), source, gbank, f);
  reconstruct_store<MODE>(s, i, e);
  if (MODE == 2) { float fa = getv(s, accepted(s), i); s.img_kinds[i] = fa <= 0 ? 0 : (fa >= 1 ? 1 : 2); }
}
// The same planes in two passes: full and empty cells (almost all) directly,
// interface cells appended to a work queue and reconstructed by eight lanes
// each (interface_gradient8, inverse_volume8), with the same arithmetic.
template <int MODE>
__global__ void __launch_bounds__(256) k_reconstruct_direct(Ctx s, int cap) {
  if (!reconstruct_gate<MODE>(s)) return;
  const int64_t i = GID(); bool defer = false;
  if (i < s.N) {
    const I3 q = coord(i, dims(s));
    const int gbank = geometry_bank(s, MODE == 2);
    const float f = MODE == 1 ? transport_f(s, q) : geometry_fraction(s, q, gbank);
    if (f <= 0 || f >= 1) reconstruct_store<MODE>(s, i, make_float4(0, 1, 0, f)); else defer = true;
    if (MODE == 2) { float fa = getv(s, accepted(s), i); s.img_kinds[i] = fa <= 0 ? 0 : (fa >= 1 ? 1 : 2); }
  }
  queue_append_block(s, defer, i, 3, cap);
}
template <int MODE>
__global__ void __launch_bounds__(256) k_reconstruct_deferred(Ctx s, int cap) {
  if (!reconstruct_gate<MODE>(s)) return;
  const QueueView qv = queue_view(s, 3, cap);
  const int64_t groups = (int64_t)gridDim.x * blockDim.x / 8;
  const int sub = (int)(threadIdx.x & 7);
  const int gbank = geometry_bank(s, MODE == 2), source = MODE == 1 ? 1 : 0;
  for (int64_t j = GID() / 8; j < qv.total; j += groups) {
    const int64_t i = qv.item(s, j); const I3 q = coord(i, dims(s));
    const float f = MODE == 1 ? transport_f(s, q) : geometry_fraction(s, q, gbank);
    const float4 e = plane8(s, q, source, gbank, f, sub);
    if (sub == 0) reconstruct_store<MODE>(s, i, e);
  }
}
template <int MODE>
static void launch_reconstruct(const Ctx& s, hipStream_t st) {
  const int cap = queue_cap(s.N, 256);
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, s, 3);
  hipLaunchKernelGGL(k_reconstruct_direct<MODE>, dim3((unsigned)((s.N + 255) / 256)), dim3(256), 0, st, s, cap);
  hipLaunchKernelGGL(k_reconstruct_deferred<MODE>, dim3(1024), dim3(256), 0, st, s, cap);
}
// pressure_pass.glslinc:131-137 (op 18)
// The exact distance in two passes: interface cells and cells with no
// opposite-side neighbour (almost all) directly; the seeded cells along the
// interface, each a 27-neighbour polygon search, with 32 lanes per cell.
__device__ __forceinline__ void write_distance(const Ctx& s, int64_t i, I3 q, float ph) {
  s.geo[i] = make_float2(ph, solid(s, q) ? -1.0f : (ph < 0 ? 1.0f : 0.0f));
}
__global__ void __launch_bounds__(256) k_distance_direct(Ctx s, int cap) {
  if (s.c[4] > .5f && s.c[5] < .5f) return;
  int64_t i = GID(); bool defer = false;
  if (i < s.N) {
    I3 q = coord(i, dims(s)); bool seed;
    float ph = distance_head(s, q, geometry_bank(s, false), seed);
    if (seed) defer = true; else write_distance(s, i, q, ph);
  }
  queue_append_block(s, defer, i, 2, cap);
}
// lane k < 27 of a 32-lane group evaluates neighbour k; lane 0 gathers the 27
// (box, candidate) pairs and runs the reference's pruned search in order
constexpr int DIST_THREADS = 64;  // threads per block of the queued distance pass
__global__ void __launch_bounds__(DIST_THREADS) k_distance_deferred(Ctx s, int cap) {
  if (s.c[4] > .5f && s.c[5] < .5f) return;
  // One 12-vertex list per thread.  polygon_distance fills up to 12 vertices
  // (a plane through cube corners meets up to 12 edges) in va; the transition
  // search clips a face square by two planes, va -> vb -> va, and a convex
  // polygon clipped by a plane gains at most one vertex (4 -> 5 -> 6), so its
  // two lists take the two halves of the same slot (the routines run one
  // after the other, never together): half the LDS of two full lists, two
  // more waves per CU.
  __shared__ float verts[VERT_FLOATS * DIST_THREADS];
  const Verts<DIST_THREADS> va{verts + threadIdx.x}, vb{verts + (VERT_FLOATS / 2) * DIST_THREADS + threadIdx.x};
  const QueueView qv = queue_view(s, 2, cap);
  const int64_t n = qv.total, groups = (int64_t)gridDim.x * blockDim.x / 32;
  const int sub = (int)(threadIdx.x & 31);
  const int gb = geometry_bank(s, false);
  for (int64_t j = GID() / 32; j < n; j += groups) {
    int64_t i = qv.item(s, j); I3 q = coord(i, dims(s));
    float box = __int_as_float(0x7f800000), m = box;
    if (sub < 27) distance_neighbour(s, q, sub, gb, box, m, va, vb);
    bool seed; float head = distance_head(s, q, gb, seed);
    float best = fabsf(head);  // 4 |h|, the search's start
    // every lane runs the same in-order search over the gathered pairs
    for (int k = 0; k < 27; k++) {
      float bk = __shfl(box, k, 32), mk = __shfl(m, k, 32);
      best = distance_step(best, bk, mk);
    }
    if (sub == 0) write_distance(s, i, q, (geometry_fraction(s, q, gb) >= 1 ? -1 : 1) * best);
  }
}
// pressure_pass.glslinc:108-115 (op 16)
// The geometry (wet flag, level set) of the cell and its six neighbours is
// read up front (index 0 for a cell outside the grid), then coefficient() and
// distance_face() are evaluated on it: the same values, without a load behind
// each face's tests.
struct CellGeo { I3 w; bool out; float2 g; };
__device__ __forceinline__ CellGeo cell_geo(const Ctx& s, I3 x) {
  CellGeo c; c.w = wrap_cell(s, x); c.out = outside(s, c.w);
  c.g = s.geo[c.out ? 0 : index3(c.w, dims(s))]; return c;
}
__device__ __forceinline__ bool geo_wet(const CellGeo& c) { return !c.out && c.g.y > 0.5f; }
__device__ __forceinline__ float geo_phi(const Ctx& s, const CellGeo& c) { return c.out ? center(s, c.w).y - eta(s, c.w) : c.g.x; }
// distance_face(s, l, r, axis) on the cells' preloaded geometry
__device__ __forceinline__ float geo_distance_face(const Ctx& s, I3 l, I3 r, int axis, const CellGeo& cl, const CellGeo& cr) {
  const float h = s.h[axis];
  if (axis == 1 && (l.y >= s.ny || r.y >= s.ny)) return .5f * h;
  const bool wl = geo_wet(cl), wr = geo_wet(cr);
  if (wl && wr) return h;
  const float pl = fabsf(geo_phi(s, cl)), pr = fabsf(geo_phi(s, cr));
  return h * gmax((wl ? pl : pr) / gmax(pl + pr, 1e-30f), 1e-6f);
}
__global__ void __launch_bounds__(256) k_coefficients(Ctx s) {
  int64_t i = GID(); if (i >= s.N) return;
  if (s.c[4] > .5f && s.c[5] < .5f) return;
  I3 q = coord(i, dims(s)); float diag = 0;
  const CellGeo cq = cell_geo(s, q); const bool w = geo_wet(cq);
  CellGeo cn[6];
#pragma unroll
  for (int f6 = 0; f6 < 6; f6++) { I3 nb = q; addc(nb, f6 >> 1, (f6 & 1) ? 1 : -1); cn[f6] = cell_geo(s, nb); }
#pragma unroll
  for (int f6 = 0; f6 < 6; f6++) {
    const int axis = f6 >> 1; I3 nb = q; addc(nb, axis, (f6 & 1) ? 1 : -1);
    float k = 0;
    // coefficient(s, q, nb, axis) with distance_face(s, q, nb, axis)
    if (w && !solid(s, nb) && !(axis == 0 && nb.x < 0)) k = 1.0f / (s.h[axis] * geo_distance_face(s, q, nb, axis, cq, cn[f6]));
    s.work[(int64_t)(6 + f6) * s.N + i] = k; diag += k;
  }
  s.work[12 * s.N + i] = w ? diag : 1.0f;
}
// pressure_pass.glslinc:104-107 (op 14, plain-F export; not in the coupled program)
__global__ void __launch_bounds__(256) k_export_plain(Ctx s) {
  int64_t i = GID(); if (i >= s.N) return;
  float f = getv(s, accepted(s), i);
  s.img_planes[i] = make_float4(0, 1, 0, f); s.img_kinds[i] = f == 0 ? 0 : (f == 1 ? 1 : 2);
}

// ------------------------------------------------------- trial control (43,1) --
// coupled_motion.glslinc:351-355 then pressure_pass.glslinc:20-24
__global__ void __launch_bounds__(256) k_retry(Ctx s) {
  float* c = s.c;
  bool retry = c[7] == 20 || (c[7] >= 10 && c[7] <= 12);
  if (retry && c[29] < 2) { c[33] = c[7]; c[34] = c[20]; c[30] = c[20] * .5f; c[29]++; c[7] = 0; c[4] = 0; c[5] = 0; }
}
__global__ void __launch_bounds__(256) k_trial_start(Ctx s) {
  float* c = s.c; float quota = *s.quota;
  c[5] = 0;
  if (c[7] > 0.5f || (quota >= 0 ? c[2] >= quota : c[1] >= -quota)) return;
  if (c[4] < 0.5f) { c[4] = 1; c[5] = 1; c[6] = 0; c[8] = 0; c[13] = 1; c[32]++; c[57] = 0; c[58] = 0; }
}
__global__ void __launch_bounds__(256) k_inject(Ctx s) { s.c[19] = 1; }  // op 15, fault injection

// Serial in-order sums of the block partials (the reference's lane-0 loops,
// e.g. pressure_pass.glslinc:80, coupled_motion.glslinc:402-408).  The
// partials are staged in LDS in chunks, each component in a contiguous run;
// lane 0 adds component CA and lane 1 component CB (CB < 0: none) in index
// order, so the two dependent chains advance in the same instructions.  The
// lanes read their runs four values at a time and load the next batch of 32
// before adding the current one, and while the first wave sums a chunk the
// second wave stages the next (double buffer).  Lane 0 returns
// (sum CA, sum CB).  Launched with WV_RED (two waves) threads.
template <int C>
__device__ __forceinline__ float component(float4 p) { return C == 0 ? p.x : (C == 1 ? p.y : (C == 2 ? p.z : p.w)); }
template <int CA, int CB>
__device__ __forceinline__ float2 serial_sums(const float4* partial, int64_t groups) {
  static_assert(WV_RED == 128, "one summing wave and one staging wave");
  constexpr int CHUNK = 2048, NC = CB >= 0 ? 2 : 1, BATCH = 7, Q = CHUNK / 4;
  __shared__ __align__(16) float stage[2][NC][CHUNK];
  const int t = threadIdx.x;
  const int chunks = (int)((groups + CHUNK - 1) / CHUNK);
  auto fill = [&](int c, int first, int stride) {
    const int64_t base = (int64_t)c * CHUNK;
    const int n = (int)(groups - base < CHUNK ? groups - base : CHUNK);
    for (int k = first; k < n; k += stride) {
      const float4 p = partial[base + k];
      stage[c & 1][0][k] = component<CA>(p);
      if (NC == 2) stage[c & 1][NC - 1][k] = component<CB < 0 ? 0 : CB>(p);
    }
  };
  float acc = 0;
  if (chunks > 0) fill(0, t, WV_RED);
  __syncthreads();
  for (int c = 0; c < chunks; c++) {
    if (t >= 64) {
      if (c + 1 < chunks) fill(c + 1, t - 64, WV_RED - 64);
    } else if (t < NC) {
      const float4* src = reinterpret_cast<const float4*>(stage[c & 1][t]);
      const int64_t left = groups - (int64_t)c * CHUNK;
      const int n = (int)(left < CHUNK ? left : CHUNK), n4 = n >> 2, nb = n4 / BATCH;
      // batches past the last full one are loaded (inside the buffer) but not added
      float4 a[BATCH], b[BATCH];
#pragma unroll
      for (int e = 0; e < BATCH; e++) a[e] = src[e];
      for (int q = 0; q < nb; q += 2) {
        const int jb = (q + 1) * BATCH < Q - BATCH ? (q + 1) * BATCH : Q - BATCH;
#pragma unroll
        for (int e = 0; e < BATCH; e++) b[e] = src[jb + e];
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the additions
#pragma unroll
        for (int e = 0; e < BATCH; e++) { acc += a[e].x; acc += a[e].y; acc += a[e].z; acc += a[e].w; }
        if (q + 1 >= nb) break;
        const int ja = (q + 2) * BATCH < Q - BATCH ? (q + 2) * BATCH : Q - BATCH;
#pragma unroll
        for (int e = 0; e < BATCH; e++) a[e] = src[ja + e];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < BATCH; e++) { acc += b[e].x; acc += b[e].y; acc += b[e].z; acc += b[e].w; }
      }
      for (int j = nb * BATCH; j < n4; j++) { const float4 v = src[j]; acc += v.x; acc += v.y; acc += v.z; acc += v.w; }
      for (int k = 4 * n4; k < n; k++) acc += stage[c & 1][t][k];
    }
    __syncthreads();
  }
  float other = __shfl(acc, 1, 64);
  return make_float2(acc, other);
}

// Up to eight in-order sums side by side, two per source array: chain 2 k
// adds component ca[k] and chain 2 k + 1 component cb[k] (cb[k] < 0: none) of
// src[k][0, groups) in index order (every src[k] readable; the caller ignores
// the chains it does not need).  The
// second wave stages the next chunk of every chain in LDS while lanes 0-7 of
// the first wave add the current one, each lane its own chain in the same
// instructions, in the loop of serial_sums.  Lane c of the first wave returns
// chain c's sum.  Launched with WV_RED (two waves) threads.
__device__ __forceinline__ float pick(float4 p, int c) { return c == 0 ? p.x : (c == 1 ? p.y : (c == 2 ? p.z : p.w)); }
template <int NS>
__device__ __forceinline__ float serial_sums_multi(const float4* const* src, const int* ca, const int* cb, int64_t groups) {
  static_assert(WV_RED == 128 && NS <= 4, "one summing wave and one staging wave, at most 8 chains");
  constexpr int CHUNK = 1024, NCH = 2 * NS, BATCH = 7, Q = CHUNK / 4;
  __shared__ __align__(16) float stage[2][NCH][CHUNK];
  const int t = threadIdx.x;
  const int chunks = (int)((groups + CHUNK - 1) / CHUNK);
  auto fill = [&](int c, int first, int stride) {
    const int64_t base = (int64_t)c * CHUNK;
    const int n = (int)(groups - base < CHUNK ? groups - base : CHUNK);
    // every source's element loaded before any is staged: NS loads in flight per step
#pragma unroll 2
    for (int e = first; e < n; e += stride) {
      float4 p[NS];
#pragma unroll
      for (int k = 0; k < NS; k++) p[k] = src[k][base + e];
#pragma unroll
      for (int k = 0; k < NS; k++) {
        stage[c & 1][2 * k][e] = pick(p[k], ca[k]);
        if (cb[k] >= 0) stage[c & 1][2 * k + 1][e] = pick(p[k], cb[k]);
      }
    }
  };
  float acc = 0;
  if (chunks > 0) fill(0, t, WV_RED);
  __syncthreads();
  for (int c = 0; c < chunks; c++) {
    if (t >= 64) {
      if (c + 1 < chunks) fill(c + 1, t - 64, WV_RED - 64);
    } else if (t < NCH) {
      const float4* run = reinterpret_cast<const float4*>(stage[c & 1][t]);
      const int64_t left = groups - (int64_t)c * CHUNK;
      const int n = (int)(left < CHUNK ? left : CHUNK), n4 = n >> 2, nb = n4 / BATCH;
      float4 a[BATCH], b[BATCH];
#pragma unroll
      for (int e = 0; e < BATCH; e++) a[e] = run[e];
      for (int q = 0; q < nb; q += 2) {
        const int jb = (q + 1) * BATCH < Q - BATCH ? (q + 1) * BATCH : Q - BATCH;
#pragma unroll
        for (int e = 0; e < BATCH; e++) b[e] = run[jb + e];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < BATCH; e++) { acc += a[e].x; acc += a[e].y; acc += a[e].z; acc += a[e].w; }
        if (q + 1 >= nb) break;
        const int ja = (q + 2) * BATCH < Q - BATCH ? (q + 2) * BATCH : Q - BATCH;
#pragma unroll
        for (int e = 0; e < BATCH; e++) a[e] = run[ja + e];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < BATCH; e++) { acc += b[e].x; acc += b[e].y; acc += b[e].z; acc += b[e].w; }
      }
      for (int j = nb * BATCH; j < n4; j++) { const float4 v = run[j]; acc += v.x; acc += v.y; acc += v.z; acc += v.w; }
      for (int k = 4 * n4; k < n; k++) acc += stage[c & 1][t][k];
    }
    __syncthreads();
  }
  return acc;
}

// ---------------------------------------------------- CFL / dt (20 + 21) -----
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_cfl(Ctx s, int) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  if (!FINAL) {
  int64_t i = GID();
  float rate = i < s.N ? courant_at(s, coord(i, dims(s)), accepted(s), true) : 0;
  float4 v = group_reduce<RED_BOUNDS>(make_float4(rate, rate, 0, 0));
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  // coupled_motion.glslinc:362-369: a max is order-free, so reduce it in parallel
  float m = 0; 
#pragma unroll 8
  for (int64_t k = threadIdx.x; k < G; k += WV_RED) m = gmax(m, s.partial[k].y);
  float4 r = group_reduce<RED_DIAG>(make_float4(m, 0, 0, 0));
  if (threadIdx.x == 0) {
    float* c = s.c; float rt = r.x;
    c[20] = gmin(s.dt_max, .45f / gmax(rt, 1e-20f)); if (c[29] > 0) c[20] = gmin(c[20], c[30]);
    c[21] = 0; c[22] = 0; c[23] = 0; c[39] = 0;
    for (int field = 64; field < 70; field++) c[field] = 0;
  }
}
// ---------------------------------------------- compression audit (45 + 46) --
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_compression(Ctx s, int) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  if (!FINAL) {
  int64_t i = GID(); float4 v = make_float4(0, 0, 0, 0);
  if (i < s.N) {
    int acc = accepted(s); float dt = timestep(s); float cell = s.h[0] * s.h[1] * s.h[2];
    I3 q = coord(i, dims(s)); float f = getv(s, acc, i), dv = 0, raw = 0;
    for (int axis = 0; axis < 3; axis++) {
      I3 r = q; addc(r, axis, 1);
      dv += (old_face(s, r, axis) - old_face(s, q, axis)) / s.h[axis];
      raw += (getv(s, acc, face_index(s, r, axis)) - getv(s, acc, face_index(s, q, axis))) / s.h[axis];
    }
    v = make_float4(f > .5f ? dt * dv * cell : 0, f > .5f && s.motion[pmask_base(s, acc) + i] < .5f ? 1 : 0, f * cell, f > .5f ? dt * raw * cell : 0);
  }
  v = group_reduce<RED_SUM>(v);
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float4 g = global_reduce<false>(s.partial, G);
  if (threadIdx.x == 0) { float* c = s.c; c[35] = g.x; c[36] = g.y; c[37] = g.z; c[38] = g.w; if (g.y > 0) { c[7] = 22; c[3]++; } }
}
// ---------------------------------------------------------- bank copy (22) --
__global__ void __launch_bounds__(256) k_bank_copy(Ctx s) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  int64_t i = GID(); int acc = accepted(s), cb = 1 - acc;
  if (i < s.stride) putv(s, cb, i, getv(s, acc, i));
  if (i < s.N) s.motion[s.N + i] = getv(s, acc, i) > .5f ? 1 : 0;
}
// -------------------------------------------------- geometric flux (24) -----
__global__ void __launch_bounds__(256) k_flux(Ctx s) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  int64_t i = GID(); if (i >= s.NF) return;
  int sweep_axis = ((int)s.c[2] + (int)s.c[21]) % 3;
  I3 q; int axis; unpack_face(s, i, q, axis);
  float amount = 0;
  if (axis == sweep_axis && !wall_face(s, q, axis)) {
    float velocity = old_face(s, q, axis), courant = face_courant(s, q, axis);
    I3 donor = q; if (velocity > 0) addc(donor, axis, -1);
    donor = wrap_cell(s, donor); float f = transport_f(s, donor), width = fabsf(courant); amount = f * width;
    if (f > 0 && f < 1 && width > 0) {
      float4 eq = donor.x < 0 ? coastal_plane(s, donor, s.c[1]) : s.plane[index3(donor, dims(s))];
      F3 low = mkf(0, 0, 0), extent = mkf(1, 1, 1); setc(extent, axis, width); if (velocity > 0) setc(low, axis, 1 - width);
      amount = width * cube_volume(mkf(eq.x * extent.x, eq.y * extent.y, eq.z * extent.z), eq.w - dot3(mkf(eq.x, eq.y, eq.z), low));
    }
    amount *= gsign(velocity);
  }
  s.motion[flux_base(s) + i] = amount;
}
// ---------------------------------------------- boundary ledgers (55 + 56) --
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_ledgers(Ctx s, int) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  if (!FINAL) {
  int64_t i = GID(); float4 v = make_float4(0, 0, 0, 0);
  if (i < s.N) {
    int axis = ((int)s.c[2] + (int)s.c[21]) % 3;
    I3 q = coord(i, dims(s));
    if (axis == 1 && q.y == s.ny - 1) { q.y++; v.x = s.motion[flux_base(s) + face_local(s, q, 1)] * s.h[0] * s.h[1] * s.h[2]; }
    if (axis == 0 && q.x == 0) { float flux = s.motion[flux_base(s) + face_local(s, q, 0)] * s.h[0] * s.h[1] * s.h[2]; v.y = gmax(flux, 0); v.z = gmax(-flux, 0); }
  }
  v = group_reduce<RED_SUM>(v);
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float4 g = global_reduce<false>(s.partial, G);
  if (threadIdx.x == 0) { s.c[39] += g.x; s.c[65] += g.y; s.c[66] += g.z; }
}
// ------------------------------------------------ VOF update (25 + 26) ------
// Ctx::vpart: the VOF update's group partials of sweep k (k = c[21], at most
// 3), then the sweep headers of the batch program
__device__ __forceinline__ int64_t red_groups(const Ctx& s) { return (s.N + WV_RED - 1) / WV_RED; }
__device__ __forceinline__ float4* vof_partials(const Ctx& s, int sweep) { return s.vpart + (int64_t)(sweep < 3 ? sweep : 3) * red_groups(s); }
__device__ __forceinline__ float4* vof_header(const Ctx& s, int j) { return s.vpart + 4 * red_groups(s) + j; }
// FINAL: the one-group finaliser, a kernel of its own so that the partial
// pass does not carry the finaliser's LDS staging (occupancy).  Its argument
// j < 0: the whole finaliser.  j = 0, 1, 2 (sweep j of the batch program):
// the bounds and the failure test only, and header j = (ran, c[21], c[32]);
// the two in-order ledger sums of the sweep (c[22], c[23] and the history's
// z, w) are added by op 12's finaliser of the same batch (k_accept), in sweep
// order.  Nothing between reads those, and they are added whether or not the
// trial goes on to op 12's acceptance, as the per-sweep finaliser adds them.
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_vof_update(Ctx s, int j) {
  const bool open = s.c[5] > .5f && s.c[7] < .5f;
  if (FINAL && j >= 0 && threadIdx.x == 0) *vof_header(s, j) = make_float4(open ? 1 : 0, s.c[21], s.c[32], 0);
  if (!open) return;
  if (!FINAL) {
  int64_t i = GID(); float4 v = make_float4(1e20f, -1e20f, 0, 0);
  int sweep = (int)s.c[21];
  if (i < s.N) {
    int axis = ((int)s.c[2] + sweep) % 3;
    I3 q = coord(i, dims(s)), r = q; addc(r, axis, 1);
    float net = s.motion[flux_base(s) + face_local(s, q, axis)] - s.motion[flux_base(s) + face_local(s, r, axis)];
    float correction = s.motion[s.N + i] * (face_courant(s, r, axis) - face_courant(s, q, axis));
    float increment = net + correction;
    float f = transport_f(s, q) + increment; float change = 0;
    if (f >= -1e-6f && f <= 1 + 1e-6f) change = gclamp(f, 0, 1) - f;
    if (sweep == 1) s.motion[i] = f + change; else putv(s, candidate(s), i, f + change);
    v = make_float4(f, f, change, fabsf(change)); if (isnan(f) || isinf(f)) { v.x = -1e30f; v.y = 1e30f; }
  }
  v = group_reduce<RED_BOUNDS>(v);
  if (threadIdx.x == 0) vof_partials(s, sweep)[blockIdx.x] = v;
  return;
  }
  const int64_t G = red_groups(s);
  const float4* part = vof_partials(s, (int)s.c[21]);
  // coupled_motion.glslinc:402-408: min/max in parallel, the two sums in index order
  float mn = 1e20f, mx = -1e20f;
#pragma unroll 8
  for (int64_t k = threadIdx.x; k < G; k += WV_RED) { mn = gmin(mn, part[k].x); mx = gmax(mx, part[k].y); }
  float4 b = group_reduce<RED_BOUNDS>(make_float4(mn, mx, 0, 0));
  float2 zw = make_float2(0, 0);
  if (j < 0) zw = serial_sums<2, 3>(part, G);
  if (threadIdx.x == 0) {
    float* c = s.c; float cell = s.h[0] * s.h[1] * s.h[2];
    uint32_t step = (uint32_t)(int64_t)c[32] - 1u, sw = (uint32_t)c[21];
    if (step < (uint32_t)s.hist_cap) {
      float* hv = (float*)&s.history[(int64_t)step * 4 + sw];
      hv[0] = b.x; hv[1] = b.y;
      if (j < 0) { hv[2] = zw.x; hv[3] = zw.y; }
    }
    if (j < 0) { c[22] += zw.x * cell; c[23] += zw.y * cell; }
    if (b.x < -1e-6f || b.y > 1 + 1e-6f) { c[7] = 10 + c[21]; c[3]++; }
    c[21]++;
  }
}
void wv_launch_vof_deferred(const Ctx& s, int sweep, hipStream_t st) {
  hipLaunchKernelGGL(k_vof_update<false>, dim3((unsigned)red_groups_host(s)), dim3(WV_RED), 0, st, s, 0);
  hipLaunchKernelGGL(k_vof_update<true>, dim3(1), dim3(WV_RED), 0, st, s, sweep);
}
// ---------------------------------------------- MacCormack (27 / 28) --------
// coupled_motion.glslinc:168-188, 409-426 (transport-cache and plain forms)
// CACHE mirrors numerics flag 4 (set for the whole run): one kernel per form
// keeps each variant's registers and code to its own path
// One face of the MacCormack pass (coupled_motion.glslinc:142-188).  With
// DEFER (the cached form's first pass) a face whose trace would need the
// general sampling path writes nothing and returns false.
template <bool CORRECT, bool CACHE, bool DEFER, bool COOP = false>
__device__ __forceinline__ bool maccormack_face(const Ctx& s, int64_t i, int sub = 0) {
  I3 q; int axis; unpack_face(s, i, q, axis);
  F3 x = mkf(((float)q.x + .5f) * s.h[0], ((float)q.y + .5f) * s.h[1], ((float)q.z + .5f) * s.h[2]);
  addc(x, axis, -(.5f * s.h[axis]));
  int acc = accepted(s); float dt = timestep(s);
  float value = 0;
  if (!wall_face(s, q, axis) && ((s.flags & 2) == 0 || s.motion[valid_base(s, acc) + i] > .5f)) {
    if (CACHE) {
      int64_t base = trace_base(s), nf = s.NF;
      if (!CORRECT) {
        Rk2Trace t = rk2_samples<DEFER, COOP>(s, x, dt, axis, false, false, mkf(0, 0, 0), sub);
        if (DEFER && t.deferred) return false;
        value = t.value;
        if (COOP && sub != 0) return true;  // lane 0 of the group writes
        s.motion[base + i] = t.l0; s.motion[base + nf + i] = t.l1;
        s.motion[base + 2 * nf + i] = t.vs && t.vm && t.ve && trace_points_clear(s, x, t.mid, t.end) ? 1 : 0;
        s.motion[base + 3 * nf + i] = t.sv.x; s.motion[base + 4 * nf + i] = t.sv.y; s.motion[base + 5 * nf + i] = t.sv.z;
      } else {
        float forward = s.motion[forward_base(s) + i];
        if (s.motion[base + 2 * nf + i] < .5f) value = forward;
        else {
          F3 sv = mkf(s.motion[base + 3 * nf + i], s.motion[base + 4 * nf + i], s.motion[base + 5 * nf + i]);
          Rk2Trace t = rk2_samples<DEFER, COOP>(s, x, -dt, axis, true, true, sv, sub);
          if (DEFER && t.deferred) return false;
          float corrected = forward + .5f * (getv(s, acc, U_(s) + i) - t.value);
          value = t.vm && t.ve && trace_points_clear(s, x, t.mid, t.end) && corrected >= s.motion[base + i] && corrected <= s.motion[base + nf + i] ? corrected : forward;
        }
      }
    } else if (!CORRECT) {
      bool ok; float a, b; value = sampled_component(s, trace_rk2(s, x, dt), axis, false, ok, a, b);
    } else {
      bool valid, vf; float a, b, d0, d1;
      float backward = sampled_component(s, trace_rk2(s, x, -dt), axis, true, valid, a, b);
      sampled_component(s, trace_rk2(s, x, dt), axis, false, vf, d0, d1);
      float forward = s.motion[forward_base(s) + i], corrected = forward + .5f * (getv(s, acc, U_(s) + i) - backward);
      value = valid && vf && trace_supported(s, x, dt) && trace_supported(s, x, -dt) && corrected >= d0 && corrected <= d1 ? corrected : forward;
    }
  }
  if (COOP && sub != 0) return true;
  if (!CORRECT) s.motion[forward_base(s) + i] = value; else putv(s, candidate(s), U_(s) + i, value);
  return true;
}
// the uncached form in one pass
template <bool CORRECT>
__global__ void __launch_bounds__(256) k_maccormack(Ctx s) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  int64_t i = GID(); if (i >= s.NF) return;
  maccormack_face<CORRECT, false, false>(s, i);
}
// The cached form in two passes.  About 8 % of the traced faces sample next to
// the bed, the ocean ghost or the periodic z edge (the general path); spread
// over the grid they put the general code into almost half of the waves.  The
// first pass traces every face whose samples are all direct reads and queues
// the others (one atomic per wave); the second traces the queued faces in
// dense waves.  Each face is computed by the same code either way.
template <bool CORRECT>
__global__ void __launch_bounds__(256) k_maccormack_direct(Ctx s, int cap) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  int64_t i = GID();
  bool defer = i < s.NF && !maccormack_face<CORRECT, true, true>(s, i);
  queue_append_block(s, defer, i, CORRECT ? 1 : 0, cap);
}
// eight lanes per queued face: the deferred traces are few and latency-bound,
// and the corners of a general sample are independent donor evaluations
template <bool CORRECT>
__global__ void __launch_bounds__(256) k_maccormack_deferred(Ctx s, int cap) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  const QueueView qv = queue_view(s, CORRECT ? 1 : 0, cap);
  const int64_t n = qv.total, groups = (int64_t)gridDim.x * blockDim.x / 8;
  const int sub = (int)(threadIdx.x & 7);
  for (int64_t j = GID() / 8; j < n; j += groups) maccormack_face<CORRECT, true, false, true>(s, qv.item(s, j), sub);
}

// The wall-search code of every in-range face at z = 0, per (component, y, x)
// (the search does not depend on z); -1 where no face is in range.
__global__ void __launch_bounds__(256) k_reflect_table(Ctx s) {
  const int64_t i = GID(), row = (int64_t)(s.nx + 1) * (s.ny + 1);
  if (i >= 3 * row) return;
  const int component = (int)(i / row), r = (int)(i - component * row);
  const int y = r / (s.nx + 1), x = r - y * (s.nx + 1);
  const I3 q = mk(x, y, 0);
  s.reflect[i] = (signed char)(face_in_bounds(s, q, component) ? reflect_search(s, q, component) : -1);
}
void wv_launch_reflect_table(const Ctx& s, hipStream_t st) {
  const int64_t n = 3 * (int64_t)(s.nx + 1) * (s.ny + 1);
  hipLaunchKernelGGL(k_reflect_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s);
}

// --------------------------------------------------- coastal relaxation -----
__device__ __forceinline__ F3 cell_velocity(const Ctx& s, int bank, I3 q) {
  F3 v;
  for (int axis = 0; axis < 3; axis++) { I3 r = q; addc(r, axis, 1); setc(v, axis, .5f * (getv(s, bank, face_index(s, q, axis)) + getv(s, bank, face_index(s, r, axis)))); }
  return v;
}
__device__ __forceinline__ bool offshore_zone(const Ctx& s, float x) { return x < s.forc[0] + s.forc[1]; }
// coastal_pass.glslinc:46-58 (ops 80 + 81)
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_relax_f(Ctx s, int) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  if (!FINAL) {
  int64_t i = GID(); float4 v = make_float4(0, 0, 0, 0);
  if (i < s.N) {
    I3 q = coord(i, dims(s));
    if (!solid(s, q)) {
      int cb = candidate(s); float time = s.c[1] + timestep(s); float cell = s.h[0] * s.h[1] * s.h[2];
      F3 cc = center(s, q);
      float old = getv(s, cb, i), rate = coastal_rate(s, cc.x);
      float target = offshore_zone(s, cc.x) ? coastal_fraction(s, q, time) : gclamp(.5f - cc.y / s.h[1], 0, 1);
      float value = old + (1 - expf(-rate * timestep(s))) * (target - old); putv(s, cb, i, value);
      F3 cv = cell_velocity(s, cb, q); float m = 1000 * (value - old) * cell;
      v = make_float4((value - old) * cell, m * cv.x, m * cv.y, m * cv.z);
    }
  }
  v = group_reduce<RED_SUM>(v);
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float4 g = global_reduce<false>(s.partial, G);
  if (threadIdx.x == 0) { s.c[64] += g.x; s.c[67] += g.y; s.c[68] += g.z; s.c[69] += g.w; }
}
// coastal_pass.glslinc:59-68 (op 82)
__global__ void __launch_bounds__(256) k_relax_faces(Ctx s) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  int64_t i = GID(); if (i >= s.NF) return;
  I3 q; int axis; unpack_face(s, i, q, axis);
  int cb = candidate(s); float time = s.c[1] + timestep(s);
  F3 position = coastal_face_position(s, q, axis); float old = getv(s, cb, U_(s) + i), value = old;
  if (!wall_face(s, q, axis)) {
    float target = offshore_zone(s, position.x) ? coastal_velocity_axis(s, position, time, axis) : 0;
    float t = 1 - expf(-coastal_rate(s, position.x) * timestep(s));
    value = old * (1 - t) + target * t;
    if (axis == 0 && q.x == 0) value = target;
  }
  putv(s, cb, U_(s) + i, value); s.motion[forward_base(s) + i] = value - old;
}
// coastal_pass.glslinc:69-77 (ops 83 + 84)
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_momentum(Ctx s, int) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  if (!FINAL) {
  int64_t i = GID(); float4 v = make_float4(0, 0, 0, 0);
  if (i < s.N) {
    I3 q = coord(i, dims(s)); F3 ch; float cell = s.h[0] * s.h[1] * s.h[2];
    for (int axis = 0; axis < 3; axis++) { I3 r = q; addc(r, axis, 1); setc(ch, axis, .5f * (s.motion[forward_base(s) + face_local(s, q, axis)] + s.motion[forward_base(s) + face_local(s, r, axis)])); }
    float m = getv(s, candidate(s), i) * 1000 * cell;
    v = make_float4(ch.x * m, ch.y * m, ch.z * m, 0);
  }
  v = group_reduce<RED_SUM>(v);
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float4 g = global_reduce<false>(s.partial, G);
  if (threadIdx.x == 0) { s.c[67] += g.x; s.c[68] += g.y; s.c[69] += g.z; }
}
// coastal_pass.glslinc:84-92 (ops 87 + 88)
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_forcing_bounds(Ctx s, int) {
  if (!(s.c[5] > .5f && s.c[7] < .5f)) return;
  if (!FINAL) {
  int64_t i = GID();
  float f = i < s.N ? getv(s, candidate(s), i) : 0;
  float4 v = (isnan(f) || isinf(f)) ? make_float4(-1e30f, 1e30f, 0, 0) : make_float4(f, f, 0, 0);
  v = group_reduce<RED_BOUNDS>(v);
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float mn = 1, mx = 0;
#pragma unroll 8
  for (int64_t k = threadIdx.x; k < G; k += WV_RED) { mn = gmin(mn, s.partial[k].x); mx = gmax(mx, s.partial[k].y); }
  float4 b = group_reduce<RED_BOUNDS>(make_float4(mn, mx, 0, 0));
  if (threadIdx.x == 0) { s.c[88] = b.x; s.c[89] = b.y; if (b.x < 0 || b.y > 1) { s.c[7] = 23; s.c[3]++; } }
}
// coastal_pass.glslinc:78-83 (op 86): eight analytic target probes
__global__ void __launch_bounds__(256) k_probes(Ctx s) {
  int i = threadIdx.x; if (i >= 8) return;
  float probe_time = s.forc[3] * 6.28318530718f / s.wave[2] * (i < 4 ? .37f : 1);
  F3 point = mkf(s.o[0] + .17f + (float)i * .23f, -s.wave[3] + (float)i * .07f, 0);
  float se, ss; coastal_surface(s, point.x, probe_time, se, ss); F3 vel = coastal_velocity(s, point, probe_time);
  s.probe[i] = make_float4(se, ss, vel.x, vel.y);
}
// coastal_pass.glslinc:102-113 (op 85): eight column-height gauges; one block
// per gauge, columns in parallel over z, the z sum in index order
__global__ void __launch_bounds__(256) k_gauges(Ctx s) {
  __shared__ float columns[1024];
  int g = blockIdx.x; int acc = accepted(s);
  float begin = s.forc[0] + s.forc[1] + .4f;
  float end = s.o[0] + (float)s.nx * s.h[0] - (s.shape == 6 ? s.forc[1] + .4f : 1.0f);
  float t = (float)g / 7; float m = begin * (1 - t) + end * t;
  int x = iclamp((int)((m - s.o[0]) / s.h[0]), 0, s.nx - 1);
  float height = 0;
  for (int z0 = 0; z0 < s.nz; z0 += 1024) {
    __syncthreads();
    for (int z = z0 + threadIdx.x; z < s.nz && z < z0 + 1024; z += blockDim.x) {
      float column = s.o[1];
      // the column's cells 24 at a time, loaded ahead of the in-order sum
      for (int y0 = 0; y0 < s.ny; y0 += 24) {
        float v[24]; bool sd[24];
#pragma unroll
        for (int u = 0; u < 24; u++) {
          const I3 q = mk(x, y0 + u < s.ny ? y0 + u : s.ny - 1, z);
          sd[u] = solid(s, q); v[u] = getv(s, acc, index3(q, dims(s)));
        }
#pragma unroll
        for (int u = 0; u < 24; u++) if (y0 + u < s.ny) column += (sd[u] ? 1 : v[u]) * s.h[1];
      }
      columns[z - z0] = column;
    }
    __syncthreads();
    if (threadIdx.x == 0) for (int z = z0; z < s.nz && z < z0 + 1024; z++) height += columns[z - z0];
  }
  if (threadIdx.x == 0) s.c[80 + g] = height / (float)s.nz;
}
// coastal_pass.glslinc:93-101 (op 89)
__global__ void __launch_bounds__(256) k_coastal_history(Ctx s) {
  float* c = s.c;
  if (c[2] <= c[91] || c[2] > s.hist_cap) return;
  int64_t row = ((int64_t)c[2] - 1) * 5;
  s.chist[row] = make_float4(c[1], c[90], c[17], c[72]);
  s.chist[row + 1] = make_float4(c[73], c[74], c[40], c[24]);
  s.chist[row + 2] = make_float4(c[80], c[81], c[82], c[83]);
  s.chist[row + 3] = make_float4(c[84], c[85], c[86], c[87]);
  s.chist[row + 4] = make_float4(c[75], c[76], c[77], c[28]);
  c[91] = c[2];
}
// coupled_motion.glslinc:338-340 (op 44)
__global__ void __launch_bounds__(256) k_history(Ctx s) {
  float* c = s.c;
  if (c[32] > 0 && c[32] <= s.hist_cap) {
    int64_t row = ((int64_t)c[32] - 1) * 4 + 3;
    s.history[row] = make_float4(c[1], c[20], c[7] > 0 ? -c[7] : (c[4] < .5f ? 1 : 0), c[2]);
  }
}

// ------------------------------------------- pressure preparation (2, 47) ---
// pressure_pass.glslinc:25-50 (op 2)
__global__ void __launch_bounds__(256) k_gravity_prep(Ctx s) {
  if (s.c[5] < 0.5f) return;
  int64_t i = GID(); if (i >= s.stride) return;
  int acc = accepted(s), cb = 1 - acc; float dt = timestep(s);
  int64_t P = P_(s), Ub = U_(s), Vb = V_(s), Wb = W_(s);
  float val = getv(s, cb, i);
  if (i < s.N) s.work[13 * s.N + i] = 0;
  if (i >= P) val = wet(s, coord(i - P, dims(s))) ? getv(s, acc, i) : 0;
  if (i >= Ub && i < P) {
    int axis = i < Vb ? 0 : (i < Wb ? 1 : 2); int64_t base = axis == 0 ? Ub : (axis == 1 ? Vb : Wb);
    I3 fn = dims(s); addc(fn, axis, 1);
    I3 r = coord(i - base, fn), l = r; addc(l, axis, -1);
    if (solid(s, l) || solid(s, r) || (!wet(s, l) && !wet(s, r))) val = 0.0f;
    else if (axis == 1) val -= 9.81f * dt * s.c[31];
  }
  putv(s, cb, i, val);
}
// coupled_motion.glslinc:319-322 (op 47)
// Only the duplicate faces of a periodic seam differ from their canonical face
// (face_index): the W faces at z = nz, and the U faces at x = nx when x is
// periodic.  One thread per duplicate face copies the canonical value, as the
// pass over all faces does for exactly these faces.
__global__ void __launch_bounds__(256) k_periodic_sync(Ctx s) {
  if (s.c[4] < .5f || s.c[7] > .5f) return;
  const int64_t t = GID(), nw = (int64_t)s.nx * s.ny, nu = s.perx > .5f ? (int64_t)s.ny * s.nz : 0;
  if (t >= nw + nu) return;
  I3 q; int axis;
  if (t < nw) { axis = 2; q = mk((int)(t % s.nx), (int)(t / s.nx), s.nz); }
  else { axis = 0; const int64_t u = t - nw; q = mk(s.nx, (int)(u % s.ny), (int)(u / s.ny)); }
  I3 n = dims(s); addc(n, axis, 1);
  const int64_t own = (axis == 0 ? U_(s) : W_(s)) + index3(q, n);
  const int64_t canonical = face_index(s, q, axis);
  if (canonical != own) { int cb = candidate(s); putv(s, cb, own, getv(s, cb, canonical)); }
}
// coupled_motion.glslinc:282-284 (ops 63, 64, 65)
__global__ void __launch_bounds__(256) k_defect_state(Ctx s) { float* c = s.c; if (c[57] == 1 && c[7] < .5f) { c[57] = 2; c[6] = 0; c[8] = 0; } }
__global__ void __launch_bounds__(256) k_defect_save(Ctx s) {
  if (!(s.c[57] == 2 && s.c[7] < .5f)) return;
  int64_t i = GID(); if (i >= s.N) return;
  int cb = candidate(s); s.work[13 * s.N + i] = getv(s, cb, P_(s) + i); putv(s, cb, P_(s) + i, 0);
}
__global__ void __launch_bounds__(256) k_defect_restore(Ctx s) {
  if (!(s.c[4] > .5f && s.c[6] > .5f && s.c[7] < .5f)) return;
  int64_t i = GID(); if (i >= s.N) return;
  int cb = candidate(s); putv(s, cb, P_(s) + i, getv(s, cb, P_(s) + i) + s.work[13 * s.N + i]);
}

// --------------------------------------------------------------- PCG ops ----
// pressure_iterations.glslinc:4-13 (op 3).  The block partials of op 3 are
// replaced by the V-cycle's output partials before op 4 reads them
// (pressure_multigrid.glslinc:230-234), so this kernel publishes none.
__global__ void __launch_bounds__(256) k_pcg_init(Ctx s) {
  if (!pressure_setup(s)) return;
  int64_t i = GID(); if (i >= s.N) return;
  I3 q = coord(i, dims(s)); float r = 0, z = 0, bv = 0, diag = 1; float dt = timestep(s);
  if (wet(s, q)) { diag = s.work[12 * s.N + i]; bv = -1000.0f / dt * divergence(s, q); r = bv - operator_at(s, q, 0); z = r / diag; }
  int64_t N = s.N;
  s.work[i] = r; s.work[N + i] = z; s.work[2 * N + i] = z; s.work[4 * N + i] = diag; s.work[5 * N + i] = bv;
}
// pressure_iterations.glslinc:20 (op 5)
__global__ void __launch_bounds__(256) k_spmv(Ctx s) {
  if (!solving(s)) return;
  int64_t i = GID(); if (i >= s.N) return;
  I3 q = coord(i, dims(s)); s.work[3 * s.N + i] = wet(s, q) ? operator_at(s, q, 1) : 0;
}
// ops 5 + 6 in one pass: A p of every cell and, per 128-cell group, the p.Ap
// partial of op 6 from the value just computed (the same groups and the same
// combine tree as k_dot_alpha's first pass, so the partials are identical)
__global__ void __launch_bounds__(WV_RED) k_spmv_dot(Ctx s) {
  if (!solving(s)) return;
  const int64_t i = GID(), N = s.N;
  float prod = 0;
  if (i < N) {
    const I3 q = coord(i, dims(s));
    const float ap = wet(s, q) ? operator_at(s, q, 1) : 0;
    s.work[3 * N + i] = ap; prod = s.work[2 * N + i] * ap;
  }
  const float4 v = group_reduce<RED_SUM>(make_float4(prod, 0, 0, 0));
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
}
// pressure_iterations.glslinc:21-22 (ops 6 + 7)
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_dot_alpha(Ctx s, int) {
  if (!solving(s)) return;
  if (!FINAL) {
  int64_t i = GID(); int64_t N = s.N;
  float4 v = i < N ? make_float4(s.work[2 * N + i] * s.work[3 * N + i], 0, 0, 0) : make_float4(0, 0, 0, 0);
  v = group_reduce<RED_SUM>(v);
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float dotp = global_reduce<false>(s.partial, G).x;
  if (threadIdx.x == 0) {
    if (dotp <= 0 || isnan(dotp) || isinf(dotp)) { s.c[7] = 1; return; }
    s.c[11] = s.c[9] / dotp;
  }
}
// pressure_iterations.glslinc:23-30 (op 8) fused with the V-cycle's load of r
// (pressure_multigrid.glslinc:191).  z = r/diag and the op-8 partials are
// superseded by the V-cycle output and are not written.
// The V-cycle that follows (op 9) starts by loading b = r, x = 0 into level 0;
// that load and, with `fuse`, the level's first Chebyshev sweep (cheb_first)
// are done here.  The update's gate solving() is the V-cycle's mg_enabled().
__global__ void __launch_bounds__(256) k_pcg_update(Ctx s, MGDev m, int fuse) {
  if (!solving(s)) return;
  int64_t i = GID(); int64_t N = s.N; if (i >= N) return;
  float alpha = s.c[11];
  float r = s.work[i] - alpha * s.work[3 * N + i];
  int cb = candidate(s);
  putv(s, cb, P_(s) + i, getv(s, cb, P_(s) + i) + alpha * s.work[2 * N + i]);
  s.work[i] = r;
  float* v = m.vec + m.lv[0].vec_off;
  v[i] = r; v[N + i] = 0;
  if (fuse) cheb_first(&v[7 * N + i], &v[2 * N + i], v[4 * N + i], r);
  else v[2 * N + i] = 0;
}
// pressure_iterations.glslinc:32 (op 10)
__global__ void __launch_bounds__(256) k_direction(Ctx s) {
  if (!solving(s)) return;
  int64_t i = GID(); int64_t N = s.N; if (i >= N) return;
  s.work[2 * N + i] = s.work[N + i] + s.c[12] * s.work[2 * N + i];
}

// --------------------------------------------------------- projection (11) --
__global__ void __launch_bounds__(256) k_project(Ctx s) {
  if (s.c[4] < 0.5f || s.c[6] < 0.5f || s.c[7] > 0.5f) return;
  int64_t i = U_(s) + GID(); int64_t P = P_(s); if (i >= P) return;
  int64_t Vb = V_(s), Wb = W_(s);
  int axis = i < Vb ? 0 : (i < Wb ? 1 : 2); int64_t base = axis == 0 ? U_(s) : (axis == 1 ? Vb : Wb);
  I3 fn = dims(s); addc(fn, axis, 1);
  I3 r = coord(i - base, fn), l = r; addc(l, axis, -1);
  int cb = candidate(s); float dt = timestep(s);
  // both cells' geometry and pressures read together with the face (index 0
  // for a cell outside the grid); press() and distance_face() on them
  const CellGeo cl = cell_geo(s, l), cr = cell_geo(s, r);
  const float face = getv(s, cb, i);
  const float p_l = getv(s, cb, P + (cl.out ? 0 : index3(cl.w, dims(s)))), p_r = getv(s, cb, P + (cr.out ? 0 : index3(cr.w, dims(s))));
  const bool wl = geo_wet(cl), wr = geo_wet(cr);
  float value = 0;
  if (!solid(s, l) && !solid(s, r) && (wl || wr))
    value = face - dt / 1000.0f * ((wr ? p_r : 0.0f) - (wl ? p_l : 0.0f)) / geo_distance_face(s, l, r, axis, cl, cr);
  if (axis == 0 && r.x == 0 && !solid(s, r)) value = face;
  putv(s, cb, i, value);
}
// ---------------------------------------------------- extension (30 - 34) --
__device__ __forceinline__ bool post_ok(const Ctx& s) { return !(s.c[4] < .5f || s.c[6] < .5f || s.c[7] > .5f); }
__global__ void __launch_bounds__(256) k_ext_seed(Ctx s) {
  if (!post_ok(s)) return;
  int64_t i = GID(); if (i >= s.NF) return;
  I3 q; int axis; unpack_face(s, i, q, axis); I3 l = q; addc(l, axis, -1);
  s.motion[source_base(s, 0) + i] = !wall_face(s, q, axis) && (wet(s, q) || wet(s, l)) ? (float)i : -1;
}
__global__ void __launch_bounds__(256) k_ext_hop(Ctx s, int rb) {
  if (!post_ok(s)) return;
  int64_t i = GID(); if (i >= s.NF) return;
  int wb = 1 - rb;
  I3 q; int axis; unpack_face(s, i, q, axis);
  float source = s.motion[source_base(s, rb) + i], best = 1e20f;
  if (!wall_face(s, q, axis)) {
    // A face that is its own source sits at distance 0; a neighbour's source
    // can only tie it from the same position, which another face of this
    // axis has only across a periodic seam (z always, x when periodic), so
    // off the seams the face keeps itself and the loop is skipped.
    const bool seam = axis == 2 ? (q.z == 0 || q.z == s.nz) : (axis == 0 && s.perx > .5f && (q.x == 0 || q.x == s.nx));
    if (source >= 0 && (int64_t)source == i && !seam) { s.motion[source_base(s, wb) + i] = source; return; }
    if (source >= 0) { I3 sq; int sa; unpack_face(s, (int64_t)source, sq, sa); F3 d = periodic_delta(s, q, sq); best = dot3(d, d); }
    // the six neighbours' sources loaded together first; the wall test (two
    // bed evaluations) only for a neighbour that has a source, in the same
    // (axis, sign) order: the same candidates as testing the wall first
    I3 nbs[6]; float css[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
      I3 nb = q; addc(nb, k >> 1, (k & 1) ? 1 : -1); nb = wrap_face(s, nb); nbs[k] = nb;
      css[k] = face_in_bounds(s, nb, axis) ? s.motion[source_base(s, rb) + face_local(s, nb, axis)] : -1.0f;
    }
    for (int k = 0; k < 6; k++) {
      const float cs = css[k];
      if (cs < 0 || wall_face(s, nbs[k], axis)) continue;
      I3 sq; int sa; unpack_face(s, (int64_t)cs, sq, sa); F3 d = periodic_delta(s, q, sq); float distance = dot3(d, d);
      if (distance < best || (distance == best && cs < source)) { best = distance; source = cs; }
    }
  } else source = -1;
  s.motion[source_base(s, wb) + i] = source;
}
__global__ void __launch_bounds__(256) k_ext_fill(Ctx s) {
  if (!post_ok(s)) return;
  int64_t i = GID(); if (i >= s.NF) return;
  I3 q; int axis; unpack_face(s, i, q, axis); I3 l = q; addc(l, axis, -1);
  const float source = s.motion[source_base(s, 1) + i];  // read with the wet flags
  if (wet(s, q) || wet(s, l)) return;
  float value = source < 0 ? 0 : extended_value(s, (int64_t)source, q);
  putv(s, candidate(s), U_(s) + i, value);
}
// coupled_motion.glslinc:323-337 (ops 48 / 49)
template <bool INITIAL>
__global__ void __launch_bounds__(256) k_masks(Ctx s) {
  if (!INITIAL && !post_ok(s)) return;
  int64_t i = GID(); int cb = candidate(s);
  if (i < s.N) {
    float flag = wet(s, coord(i, dims(s))) ? 1 : 0;
    if (INITIAL) { s.motion[pmask_base(s, 0) + i] = flag; s.motion[pmask_base(s, 1) + i] = flag; }
    else s.motion[pmask_base(s, cb) + i] = flag;
  }
  if (i < s.NF) {
    I3 q; int axis; unpack_face(s, i, q, axis); I3 l = q; addc(l, axis, -1);
    float valid = INITIAL ? (!wall_face(s, q, axis) && (wet(s, q) || wet(s, l) || center(s, q).y < eta(s, q) + 3 * s.h[1]) ? 1 : 0)
                          : (s.motion[source_base(s, 1) + i] >= 0 ? 1 : 0);
    if (INITIAL) { s.motion[valid_base(s, 0) + i] = valid; s.motion[valid_base(s, 1) + i] = valid; }
    else s.motion[valid_base(s, cb) + i] = valid;
  }
}
// --------------------------------------------- post-projection CFL (35+36) --
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_post_courant(Ctx s, int) {
  if (!post_ok(s)) return;
  if (!FINAL) {
  int64_t i = GID();
  float rate = i < s.N ? courant_at(s, coord(i, dims(s)), candidate(s), false) * timestep(s) : 0;
  float4 v = group_reduce<RED_BOUNDS>(make_float4(rate, rate, 0, 0));
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float m = 0;
#pragma unroll 8
  for (int64_t k = threadIdx.x; k < G; k += WV_RED) m = gmax(m, s.partial[k].y);
  float4 r = group_reduce<RED_DIAG>(make_float4(m, 0, 0, 0));
  if (threadIdx.x == 0) s.c[27] = r.x;
}
// ------------------------------------------------ true residual (62 + 66) --
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_true_residual(Ctx s, int) {
  if (!post_ok(s)) return;
  if (!FINAL) {
  int64_t i = GID(); float r = 0;
  if (i < s.N) { I3 q = coord(i, dims(s)); if (wet(s, q)) r = 1000 / timestep(s) * divergence(s, q); }
  float4 v = group_reduce<RED_SUM>(make_float4(r * r, 0, 0, 0));
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float r2 = global_reduce<false>(s.partial, G).x;
  if (threadIdx.x == 0) {
    float* c = s.c;
    c[60] = sqrtf(r2 / gmax(c[59], 1e-30f)); c[62] = pressure_scaled_norm(s, c[59]); c[63] = pressure_scaled_norm(s, r2);
    c[61] = pressure_norm_converged(s, r2, c[59]) ? 1 : 0;
  }
}
// -------------------------------------- diagnostics + acceptance (12 + 13) --
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_accept(Ctx s, int vof_ledgers) {
  const bool ok = post_ok(s);
  if (!FINAL) {
  if (!ok) return;
  int64_t i = GID(); float4 value = make_float4(0, 0, 0, 0); float dt = timestep(s);
  if (i < s.N) {
    int cb = candidate(s); I3 q = coord(i, dims(s)); float speed = 0;
    for (int axis = 0; axis < 3; axis++) {
      I3 r = q; addc(r, axis, 1);
      speed = gmax(speed, gmax(fabsf(getv(s, cb, face_index(s, q, axis))), fabsf(getv(s, cb, face_index(s, r, axis)))));
    }
    bool w = wet(s, q);
    float div = w ? dt * fabsf(divergence(s, q)) : 0;
    float err = w ? fabsf(press(s, q) - 9810.0f * (eta(s, q) - center(s, q).y)) : 0;
    value = make_float4(div, speed, err, getv(s, cb, i) * s.h[0] * s.h[1] * s.h[2]);
    if (isnan(value.x) || isnan(value.y) || isnan(value.z) || isnan(value.w) || isinf(value.x) || isinf(value.y) || isinf(value.z) || isinf(value.w))
      value = make_float4(1e30f, 1e30f, 1e30f, 1e30f);
  }
  value = group_reduce<RED_DIAG>(value);
  if (threadIdx.x == 0) s.partial[blockIdx.x] = value;
  return;
  }
  const int64_t G = red_groups(s);
  // vof_ledgers (the batch program): the VOF sweeps' in-order ledger sums
  // left by their finalisers (k_vof_update), added first, in sweep order,
  // whether or not the trial is accepted
  float4 h[3];
  bool pending = false;
#pragma unroll
  for (int j = 0; j < 3; j++) { h[j] = vof_ledgers ? *vof_header(s, j) : make_float4(0, 0, 0, 0); pending = pending || h[j].x > .5f; }
  if (!ok && !pending) return;
  // pressure_pass.glslinc:78-103: maxima in parallel, the volume in index order
  float4 mx = make_float4(0, 0, 0, 0);
  if (ok) {
#pragma unroll 8
    for (int64_t k = threadIdx.x; k < G; k += WV_RED) { float4 p = s.partial[k]; mx.x = gmax(mx.x, p.x); mx.y = gmax(mx.y, p.y); mx.z = gmax(mx.z, p.z); }
  }
  mx = group_reduce<RED_DIAG>(mx);
  // chains 2 j, 2 j + 1: sweep j's z and w ledgers; chain 6: the volume
  const float4* src[4];
#pragma unroll
  for (int j = 0; j < 3; j++) src[j] = vof_partials(s, h[j].x > .5f ? (int)h[j].y : 0);
  src[3] = s.partial;
  const int ca[4] = {2, 2, 2, 3}, cb[4] = {3, 3, 3, -1};
  const float chain = serial_sums_multi<4>(src, ca, cb, G);
  float sums[7];
#pragma unroll
  for (int k = 0; k < 7; k++) sums[k] = __shfl(chain, k, 64);
  if (threadIdx.x != 0) return;
  {
    float* c = s.c; const float cell = s.h[0] * s.h[1] * s.h[2];
    for (int j = 0; j < 3; j++) {
      if (h[j].x < .5f) continue;
      const uint32_t step = (uint32_t)(int64_t)h[j].z - 1u, sw = (uint32_t)h[j].y;
      if (step < (uint32_t)s.hist_cap) { float* hv = (float*)&s.history[(int64_t)step * 4 + sw]; hv[2] = sums[2 * j]; hv[3] = sums[2 * j + 1]; }
      c[22] += sums[2 * j] * cell; c[23] += sums[2 * j + 1] * cell;
      *vof_header(s, j) = make_float4(0, 0, 0, 0);
    }
  }
  if (!ok) return;
  const float vol = sums[6];
  float* c = s.c; float4 v = make_float4(mx.x, mx.y, mx.z, vol); float dt = timestep(s);
  c[14] = v.x; c[15] = v.y; c[16] = v.z; c[17] = v.w;
  if ((v.x >= 1e-4f || c[61] < .5f) && v.y < 1e20f && c[19] < .5f && c[58] < 4) { c[57] = 1; c[58]++; c[6] = 0; return; }
  if (c[61] < .5f) { c[3]++; c[7] = 4; return; }
  if (v.x >= 1e-4f || v.y >= 1e20f || c[19] > 0.5f) { c[3]++; c[7] = 4; return; }
  c[28] = v.w + c[40] + c[39] - c[26] - c[24] - c[22];
  c[28] += c[74] + c[66] - c[73] - c[65] - c[72] - c[64];
  if (fabsf(c[28]) > 1e-3f * gmax(c[26], 1e-20f)) { c[3]++; c[7] = 21; return; }
  if (c[27] > .450001f) { c[3]++; c[7] = 20; return; }
  c[24] += c[22]; c[25] += c[23]; c[40] += c[39]; c[29] = 0; c[30] = 0;
  for (int field = 0; field < 6; field++) c[72 + field] += c[64 + field];
  c[90] = dt;
  c[0] = (float)candidate(s); c[1] += dt; c[2]++; c[4] = 0;
}
void wv_launch_accept(const Ctx& s, bool vof_ledgers, hipStream_t st) {
  hipLaunchKernelGGL(k_accept<false>, dim3((unsigned)red_groups_host(s)), dim3(WV_RED), 0, st, s, 0);
  hipLaunchKernelGGL(k_accept<true>, dim3(1), dim3(WV_RED), 0, st, s, vof_ledgers ? 1 : 0);
}
// ------------------------------------------------ initial volume (41 + 42) --
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_volume(Ctx s, int) {
  if (!FINAL) {
  int64_t i = GID();
  float4 v = i < s.N ? make_float4(getv(s, accepted(s), i) * s.h[0] * s.h[1] * s.h[2], 0, 0, 0) : make_float4(0, 0, 0, 0);
  v = group_reduce<RED_SUM>(v);
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float4 g = global_reduce<false>(s.partial, G);
  if (threadIdx.x == 0) s.c[26] = g.x;
}

// ---------------------------------------- pressure-section timing (telemetry) --
// The in-loop pressure time of an accepted step: device wall-clock stamps
// around the batch's pressure section (hierarchy audit and rebuild, PCG set-up,
// iterations, projection and the defect-correction restore), summed over the
// batches that section ran in -- a trial's main solve and its defect
// corrections -- and recorded when the trial is accepted.  Not a reference op:
// the reference's host measures whole submissions only
// (water/work_scheduler.gd:18-40).  Single-thread kernels; they read c[] and
// write only tstate / step_times.
__global__ void __launch_bounds__(64) k_press_begin(Ctx s) {
  if (threadIdx.x != 0) return;
  const float* c = s.c;
  unsigned long long* t = s.tstate;
  if ((unsigned long long)c[32] != t[5]) { t[2] = 0; t[6] = 0; t[5] = (unsigned long long)c[32]; }
  // the section does work when a trial is open, not failed, and its solve
  // (main or defect correction) has not converged yet
  const bool active = c[4] > .5f && c[7] < .5f && (c[6] < .5f || c[5] > .5f || c[57] == 2);
  t[1] = active ? 1 : 0;
  t[0] = wall_clock64();
}
__global__ void __launch_bounds__(64) k_press_end(Ctx s) {
  if (threadIdx.x != 0) return;
  unsigned long long* t = s.tstate;
  if (!t[1]) return;
  const unsigned long long dt = (unsigned long long)((double)(wall_clock64() - t[0]) * (double)s.tick_ns);
  t[2] += dt; t[3] += dt; t[6]++; t[1] = 0;
}
__global__ void __launch_bounds__(64) k_step_record(Ctx s) {
  if (threadIdx.x != 0) return;
  const float* c = s.c;
  unsigned long long* t = s.tstate;
  const unsigned long long steps = (unsigned long long)c[2];
  if (steps == t[4]) return;
  t[4] = steps;
  if (steps >= 1 && steps <= (unsigned long long)s.hist_cap) {
    unsigned long long* row = s.step_times + 4 * (steps - 1);
    row[0] = t[2]; row[1] = t[3]; row[2] = (unsigned long long)c[32]; row[3] = t[6];
  }
  t[3] = 0;
}
void wv_launch_press_begin(const Ctx& s, hipStream_t st) { hipLaunchKernelGGL(k_press_begin, dim3(1), dim3(64), 0, st, s); }
void wv_launch_press_end(const Ctx& s, hipStream_t st) { hipLaunchKernelGGL(k_press_end, dim3(1), dim3(64), 0, st, s); }
void wv_launch_step_record(const Ctx& s, hipStream_t st) { hipLaunchKernelGGL(k_step_record, dim3(1), dim3(64), 0, st, s); }

// ------------------------------------------------------------- launchers ----
static inline unsigned blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
#define L256(k, n) hipLaunchKernelGGL(k, dim3(blocks(n, 256)), dim3(256), 0, st, s)
// a reduction op: per-group partials over all cells, then one group finalises
// (a separate launch: the kernel boundary publishes the partials to every XCD)
#define LRED(k, id) do { hipLaunchKernelGGL(k, dim3(blocks(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0); \
                         hipLaunchKernelGGL(k, dim3(1), dim3(WV_RED), 0, st, s, 1); } while (0)
#define LRED2(k) do { hipLaunchKernelGGL(k<false>, dim3(blocks(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0); \
                      hipLaunchKernelGGL(k<true>, dim3(1), dim3(WV_RED), 0, st, s, 1); } while (0)
#define L1(k) hipLaunchKernelGGL(k, dim3(1), dim3(64), 0, st, s)

int wv_launch_op(const Ctx& s, int op, hipStream_t st) {
  switch (op) {
  case 0: L256(k_op0, s.stride); break;
  case 40: L256(k_op40, s.stride); break;
  case 90: L256(k_perturb, s.N); break;
  case 1: hipLaunchKernelGGL(k_trial_start, dim3(1), dim3(1), 0, st, s); break;
  case 43: hipLaunchKernelGGL(k_retry, dim3(1), dim3(1), 0, st, s); break;
  case 15: hipLaunchKernelGGL(k_inject, dim3(1), dim3(1), 0, st, s); break;
  case 2: L256(k_gravity_prep, s.stride); break;
  case 3: L256(k_pcg_init, s.N); break;
  case 5: L256(k_spmv, s.N); break;
  case 6: LRED2(k_dot_alpha); break;
  case 10: L256(k_direction, s.N); break;
  case 11: L256(k_project, s.NF); break;
  case 12: wv_launch_accept(s, false, st); break;
  case 14: L256(k_export_plain, s.N); break;
  case 16: L256(k_coefficients, s.N); break;
  case 17: launch_reconstruct<0>(s, st); break;
  case 18:
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, s, 2);
    hipLaunchKernelGGL(k_distance_direct, dim3(blocks(s.N, 64)), dim3(64), 0, st, s, queue_cap(s.N, 64));  // one wave per block
    hipLaunchKernelGGL(k_distance_deferred, dim3(QUEUE_BLOCKS * 256 / DIST_THREADS), dim3(DIST_THREADS), 0, st, s, queue_cap(s.N, 64));
    break;
  case 19: launch_reconstruct<2>(s, st); break;
  case 20: LRED2(k_cfl); break;
  case 22: L256(k_bank_copy, s.stride); break;
  case 23: launch_reconstruct<1>(s, st); break;
  case 24: L256(k_flux, s.NF); break;
  case 25:
    hipLaunchKernelGGL(k_vof_update<false>, dim3(blocks(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0);
    hipLaunchKernelGGL(k_vof_update<true>, dim3(1), dim3(WV_RED), 0, st, s, -1);
    break;
  case 27: case 28: {
    const bool corr = op == 28;
    if (!(s.flags & 4)) { if (corr) L256(k_maccormack<true>, s.NF); else L256(k_maccormack<false>, s.NF); break; }
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, s, corr ? 1 : 0);
    // one wave per block: waves whose faces take long paths hold no other wave's slot
    const int cap = queue_cap(s.NF, 64);
    if (corr) hipLaunchKernelGGL(k_maccormack_direct<true>, dim3(blocks(s.NF, 64)), dim3(64), 0, st, s, cap);
    else hipLaunchKernelGGL(k_maccormack_direct<false>, dim3(blocks(s.NF, 64)), dim3(64), 0, st, s, cap);
    if (corr) hipLaunchKernelGGL(k_maccormack_deferred<true>, dim3(QUEUE_BLOCKS * 4), dim3(64), 0, st, s, cap);  // one wave per block
    else hipLaunchKernelGGL(k_maccormack_deferred<false>, dim3(QUEUE_BLOCKS * 4), dim3(64), 0, st, s, cap);
    break;
  }
  case 30: L256(k_ext_seed, s.NF); break;
  case 31: case 32: case 33: hipLaunchKernelGGL(k_ext_hop, dim3(blocks(s.NF, 64)), dim3(64), 0, st, s, (op - 31) % 2); break;  // one wave per block
  case 34: L256(k_ext_fill, s.NF); break;
  case 35: LRED2(k_post_courant); break;
  case 41: LRED2(k_volume); break;
  case 44: hipLaunchKernelGGL(k_history, dim3(1), dim3(1), 0, st, s); break;
  case 45: LRED2(k_compression); break;
  case 47: L256(k_periodic_sync, (int64_t)s.nx * s.ny + (int64_t)s.ny * s.nz); break;  // the seams' duplicate faces
  case 48: L256(k_masks<false>, s.NF > s.N ? s.NF : s.N); break;
  case 49: L256(k_masks<true>, s.NF > s.N ? s.NF : s.N); break;
  case 55: LRED2(k_ledgers); break;
  case 62: LRED2(k_true_residual); break;
  case 63: hipLaunchKernelGGL(k_defect_state, dim3(1), dim3(1), 0, st, s); break;
  case 64: L256(k_defect_save, s.N); break;
  case 65: L256(k_defect_restore, s.N); break;
  case 80: LRED2(k_relax_f); break;
  case 82: hipLaunchKernelGGL(k_relax_faces, dim3(blocks(s.NF, 64)), dim3(64), 0, st, s); break;  // one wave per block
  case 83: LRED2(k_momentum); break;
  case 85: hipLaunchKernelGGL(k_gauges, dim3(8), dim3(256), 0, st, s); break;
  case 86: hipLaunchKernelGGL(k_probes, dim3(1), dim3(64), 0, st, s); break;
  case 87: LRED2(k_forcing_bounds); break;
  case 89: hipLaunchKernelGGL(k_coastal_history, dim3(1), dim3(1), 0, st, s); break;
  // finalisers fused into the kernel of the preceding op, and PCG reductions
  // fused into the V-cycle output kernel (op 4 and op 9)
  case 7: case 13: case 21: case 26: case 36: case 42: case 46: case 56: case 66: case 81: case 84: case 88:
  case 4: case 9: break;
  case 8: return -2;  // needs the MG level-0 vectors: launched through wv_launch_pcg_update
  default: return -1;
  }
  return 0;
}
// ops 5, 6 and 7 of one PCG iteration: A p with the p.Ap partials, then alpha
void wv_launch_spmv_dot(const Ctx& s, hipStream_t st) {
  hipLaunchKernelGGL(k_spmv_dot, dim3(blocks(s.N, WV_RED)), dim3(WV_RED), 0, st, s);
  hipLaunchKernelGGL(k_dot_alpha<true>, dim3(1), dim3(WV_RED), 0, st, s, 1);
}
void wv_launch_pcg_update(const Ctx& s, const MGDev& m, hipStream_t st) {
  hipLaunchKernelGGL(k_pcg_update, dim3(blocks(s.N, 256)), dim3(256), 0, st, s, m, (int)wv_mg_first_sweep_fused(m));
}
// masked 7-point SpMV launched on its own for live timing (writes only Ap)
__global__ void __launch_bounds__(256) k_spmv_timed(Ctx s) {
  int64_t i = GID(); if (i >= s.N) return;
  I3 q = coord(i, dims(s)); s.work[3 * s.N + i] = wet(s, q) ? operator_at(s, q, 1) : 0;
}
void wv_launch_spmv_timed(const Ctx& s, hipStream_t st) { L256(k_spmv_timed, s.N); }
