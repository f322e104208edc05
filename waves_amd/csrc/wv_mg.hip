// Multigrid preconditioner of the pressure solve on gfx950.
//
// The hierarchy, transfers, smoother and coarse solve follow
// water/pressure_multigrid.gd and water/shaders/compute/pressure_multigrid.glslinc,
// pressure_mg_coarse.glslinc, pressure_hierarchy_state.glsl:
//   * level 0 applies the lagged snapshot of the fine 7-point coefficients;
//   * coarse levels hold Galerkin P^T A P rows (k-major, FP32, column implicit in
//     the stencil offset);
//   * 3 + 3 sweeps of 4th-kind Chebyshev with l1-Jacobi scaling per level;
//   * trilinear transfers to levels 1 and 2, aggregation below;
//   * dense Cholesky (stored as L^-1) on the last level (<= 64 nodes).
// The reference assembles P^T A P by scattering float CAS atomics
// (pressure_multigrid.glslinc:128-175), whose order is nondeterministic.  Here
// every coarse row gathers its own entries (one thread per coarse row, the
// accumulator in LDS), then the lower triangle is copied from the upper one so
// the coarse operator is exactly symmetric, as the reference's canonical upper
// triangle storage is.  The build is deterministic.
#include "wv_mgdev.h"

// ------------------------------------------------------------ V-cycle -------
// Levels whose pre-smoothing sweep 1 is done by the kernel that loads their b
// (cheb_first): every level that is smoothed, i.e. all but the last.
static inline bool first_sweep_fused(const MGDev& m, int l) { return l < m.L - 1; }
bool wv_mg_first_sweep_fused(const MGDev& m) { return first_sweep_fused(m, 0); }

// pressure_multigrid.glslinc:191 (op 5): b = r, x1 = x2 = 0 on level 0, and
// with `fuse` the level's first Chebyshev sweep from that x
__global__ void __launch_bounds__(256) k_mg_load(Ctx s, MGDev m, int phase, int fuse) {
  if (!mg_enabled(s, phase)) return;
  int64_t i = GID(); if (i >= m.lv[0].count) return;
  const float b = s.work[i];
  m.vec[at(m, 0, 0, i)] = b; m.vec[at(m, 0, 1, i)] = 0;
  if (fuse) cheb_first(&m.vec[at(m, 0, 7, i)], &m.vec[at(m, 0, 2, i)], m.vec[at(m, 0, 4, i)], b);
  else m.vec[at(m, 0, 2, i)] = 0;
}
// pressure_multigrid.glslinc:192-196: 4th-kind Chebyshev sweep k, l1-Jacobi scaled
template <bool H>
__global__ void __launch_bounds__(256) k_mg_cheb(Ctx s, MGDev m, int l, int source, int k, int phase) {
  if (!mg_enabled(s, phase)) return;
  int64_t i = GID(); if (i >= m.lv[l].count) return;
  float r = m.vec[at(m, l, 0, i)] - mapply0<H>(m, source, (int)i);
  float d = cheb_dir(k, m.vec[at(m, l, 7, i)], m.vec[at(m, l, 4, i)], r);
  m.vec[at(m, l, 7, i)] = d; m.vec[at(m, l, 3 - source, i)] = m.vec[at(m, l, source, i)] + d;
}
// ---- coarse levels (l >= 1): G lanes per row --------------------------------
// Small levels cannot fill the chip with one thread per row, and a 125-wide row
// walked by one thread is a chain of dependent loads.  Here G lanes share a row:
// lane t takes stencil entries t, t+G, ..., then a butterfly over the G lanes
// sums the partials (a fixed order, so results are deterministic).  Interior
// rows address neighbours by a precomputed linear delta; rows near an edge go
// through the periodic/clamped wrap of pressure_multigrid.glslinc:36-41.
struct RowCtx { int row, lane, G; bool valid; };
__device__ __forceinline__ RowCtx row_ctx(int G, int N) {
  // G is a power of two: shift and mask instead of a 64-bit division
  RowCtx r; int64_t gid = GID(); int lg = __ffs(G) - 1;
  r.G = G; r.row = (int)(gid >> lg); r.lane = (int)(gid & (G - 1)); r.valid = r.row < N; return r;
}
__device__ __forceinline__ float group_sum(float v, int G) {
  for (int off = G >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// When G lanes share a row, the lanes' products go to the row's LDS slice
// (entry k at prod[k * R + r], R rows per block) and lane 0 of the row adds
// them in k order -- the order of the reference's loop
// (pressure_multigrid.glslinc:98-104) -- so the row sum is bit-identical to
// one thread walking the row.  A G > 1 call must be made by every thread of
// the block (it contains barriers); G is uniform over a launch.
#define MG_ROW_PROD_MAX 64   // narrow rows (k_mg_coarse): at most 64 entries
// sum_k A[row, k] v_field[col_k]; lane 0 of the row returns the row total in
// stencil order (every lane when G == 1)
template <bool H, int WC = 0>
__device__ __forceinline__ float coarse_apply(const MGDev& m, int l, int field, const RowCtx& rc, const int* sh_off, const int* sh_delta, float* prod) {
  const MGLevel& L = m.lv[l]; int N = L.count, W = L.width;
  const int R = (int)blockDim.x / rc.G, r = (int)threadIdx.x / rc.G;
  float sum = 0;
  auto take = [&](int k, float p) { if (rc.G == 1) sum += p; else prod[k * R + r] = p; };
  if (rc.valid) {
    I3 q = coord(rc.row, mdims(m, l));
    bool interior = q.x >= L.margin[0] && q.x < L.n[0] - L.margin[0] && q.y >= L.margin[1] && q.y < L.n[1] - L.margin[1] &&
                    q.z >= L.margin[2] && q.z < L.n[2] - L.margin[2];
    const float* row = m.rows + L.row_off + rc.row;
    const uint16_t* row16 = m.rows16 + L.row_off + rc.row;
    const float* v = m.vec + L.vec_off + (int64_t)field * N;
    auto coef = [&](int k) { return H ? bf16_to_f32(row16[(int64_t)k * N]) : row[(int64_t)k * N]; };
    // unrolled so several coefficient and neighbour loads are in flight
    if (L.wrap1) {
      // every row on one path: row + delta, one period added or removed where
      // the offset crosses the periodic z (x) edge.  Entries towards the
      // non-periodic edges hold a zero coefficient (no coarse node there), so
      // their index only has to stay inside the level.
      const int nz = L.n[2], nx = L.n[0];
      auto entry = [&](int k) {
        int j = rc.row + sh_delta[k];
        int zz = q.z + sh_off[3 * k + 2];
        j += zz < 0 ? N : (zz >= nz ? -N : 0);
        if (m.perx) { int xx = q.x + sh_off[3 * k]; j += xx < 0 ? nx : (xx >= nx ? -nx : 0); }
        j = min(max(j, 0), N - 1);
        return coef(k) * v[j];
      };
      if (WC > 0) {
        // the row width known at compile time: every entry's coefficient and
        // neighbour load is issued before the first addition (one memory
        // round trip per row), the additions still in k order
        float pr[WC > 0 ? WC : 1];
#pragma unroll
        for (int k = 0; k < WC; k++) pr[k] = entry(k);
#pragma unroll
        for (int k = 0; k < WC; k++) sum += pr[k];
      } else if (rc.G == 1) {
        // one row per thread (the large coarse levels): a unit stride lets the
        // compiler keep more of the row's loads in flight
#pragma unroll 13
        for (int k = 0; k < W; k++) sum += entry(k);
      } else {
#pragma unroll 8
        for (int k = rc.lane; k < W; k += rc.G) take(k, entry(k));
      }
    } else if (interior) {
#pragma unroll 8
      for (int k = rc.lane; k < W; k += rc.G) take(k, coef(k) * v[rc.row + sh_delta[k]]);
    } else {
#pragma unroll 4
      for (int k = rc.lane; k < W; k += rc.G) {
        int j = mwrap(m, l, mk(q.x + sh_off[3 * k], q.y + sh_off[3 * k + 1], q.z + sh_off[3 * k + 2]));
        take(k, coef(k) * v[j >= 0 ? j : 0]);
      }
    }
  }
  if (rc.G == 1) return sum;
  __syncthreads();
  if (rc.valid && rc.lane == 0) for (int k = 0; k < W; k++) sum += prod[k * R + r];
  __syncthreads();
  return sum;
}
__device__ __forceinline__ void stage_stencil(const MGDev& m, int l, int* sh_off, int* sh_delta) {
  const MGLevel& L = m.lv[l];
  for (int t = threadIdx.x; t < L.width; t += blockDim.x) {
    const int* o = m.offsets + 3 * (L.off_start + t);
    sh_off[3 * t] = o[0]; sh_off[3 * t + 1] = o[1]; sh_off[3 * t + 2] = o[2]; sh_delta[t] = m.deltas[L.off_start + t];
  }
  __syncthreads();
}
// The row epilogue (OP 0: Chebyshev sweep k from `source`; 1: residual of
// `source`; 2: Jacobi sweep from `source`; inactive rows are the identity,
// pressure_multigrid.glslinc:91) in two parts: its inputs read ahead of the
// row sum (the sweep writes them only in the row's own epilogue), then the
// update
struct EpiIn { float xs, act, b, dinv, dprev; };
template <int OP>
__device__ __forceinline__ EpiIn epilogue_load(const MGDev& m, int l, int source, int k, int i) {
  EpiIn e{0, 0, 0, 0, 0};
  e.xs = m.vec[at(m, l, source, i)]; e.act = m.vec[at(m, l, 3, i)]; e.b = m.vec[at(m, l, 0, i)];
  if (OP != 1) e.dinv = m.vec[at(m, l, 4, i)];
  if (OP == 0 && k != 1) e.dprev = m.vec[at(m, l, 7, i)];
  return e;
}
template <int OP>
__device__ __forceinline__ void epilogue_store(const MGDev& m, int l, int source, int k, int i, float ax, const EpiIn& e) {
  if (!(e.act > .5f)) ax = e.xs;  // mactive(m, l, i)
  float r = e.b - ax;
  if (OP == 1) { m.vec[at(m, l, 6, i)] = r; return; }
  if (OP == 2) { m.vec[at(m, l, 3 - source, i)] = e.xs + e.dinv * r; return; }
  float d = cheb_dir(k, e.dprev, e.dinv, r);
  m.vec[at(m, l, 7, i)] = d; m.vec[at(m, l, 3 - source, i)] = e.xs + d;
}
template <int OP, bool H, int WC>
__device__ __forceinline__ void coarse_row(const MGDev& m, int l, int source, int k, const RowCtx& rc, const int* sh_off, const int* sh_delta, float* prod) {
  const bool owner = rc.valid && rc.lane == 0;
  EpiIn e{0, 0, 0, 0, 0};
  if (owner) e = epilogue_load<OP>(m, l, source, k, rc.row);
  float ax = coarse_apply<H, WC>(m, l, source, rc, sh_off, sh_delta, prod);
  if (!owner) return;
  epilogue_store<OP>(m, l, source, k, rc.row, ax, e);
}
// Wide rows: the WIDE_WAVES waves of a block share 64 consecutive rows, wave
// w taking entries w, w + WIDE_WAVES, ... with every load 64 rows wide (256
// contiguous bytes); each product goes to LDS (entry e of the block's rows at
// prod[e][lane]) and wave 0 adds a row's products in entry order, the order
// of the reference's loop (pressure_multigrid.glslinc:98-104).
#define MG_WIDE_MAX 128      // entries of a wide row (the 125-point Galerkin stencil)
template <int OP, bool H, int WIDE_WAVES>
__global__ void __launch_bounds__(64 * WIDE_WAVES) k_mg_coarse_wide(Ctx s, MGDev m, int l, int source, int k, int phase) {
  if (!mg_enabled(s, phase)) return;
  __shared__ float prod[MG_WIDE_MAX][64];
  const MGLevel& L = m.lv[l]; const int N = L.count, W = L.width;
  // a wave's entries are the same for all its lanes: their offsets come
  // through scalar loads, no staging in LDS
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63, row = blockIdx.x * 64 + lane;
  const int* sh_off = m.offsets + 3 * L.off_start; const int* sh_delta = m.deltas + L.off_start;
  // wave 0 reads its rows' epilogue inputs before the join
  EpiIn e{0, 0, 0, 0, 0};
  if (wave == 0 && row < N) e = epilogue_load<OP>(m, l, source, k, row);
  if (row < N) {
    I3 q = coord(row, mdims(m, l));
    const float* A = m.rows + L.row_off + row;
    const uint16_t* A16 = m.rows16 + L.row_off + row;
    const float* v = m.vec + L.vec_off + (int64_t)source * N;
    if (L.wrap1) {
      const int nz = L.n[2], nx = L.n[0];
#pragma unroll 8
      for (int e = wave; e < W; e += WIDE_WAVES) {
        int j = row + sh_delta[e], zz = q.z + sh_off[3 * e + 2];
        j += zz < 0 ? N : (zz >= nz ? -N : 0);
        if (m.perx) { int xx = q.x + sh_off[3 * e]; j += xx < 0 ? nx : (xx >= nx ? -nx : 0); }
        j = min(max(j, 0), N - 1);
        prod[e][lane] = (H ? bf16_to_f32(A16[(int64_t)e * N]) : A[(int64_t)e * N]) * v[j];
      }
    } else {
      for (int e = wave; e < W; e += WIDE_WAVES) {
        int j = mwrap(m, l, mk(q.x + sh_off[3 * e], q.y + sh_off[3 * e + 1], q.z + sh_off[3 * e + 2]));
        prod[e][lane] = (H ? bf16_to_f32(A16[(int64_t)e * N]) : A[(int64_t)e * N]) * v[j >= 0 ? j : 0];
      }
    }
  }
  __syncthreads();
  if (wave != 0 || row >= N) return;
  float ax = 0;
#pragma unroll 8
  for (int e = 0; e < W; e++) ax += prod[e][lane];
  epilogue_store<OP>(m, l, source, k, row, ax, e);
}
// WC > 0: the level's width, one lane per row and the wrap-once addressing,
// known at compile time (the 39-point level of the 310x48x80 hierarchy)
template <int OP, bool H, int WC = 0>
__global__ void __launch_bounds__(256) k_mg_coarse(Ctx s, MGDev m, int l, int source, int k, int phase) {
  if (!mg_enabled(s, phase)) return;
  __shared__ int sh_off[3 * 125], sh_delta[125];
  extern __shared__ float prod[];   // (256 / G) rows x W entries when G > 1
  stage_stencil(m, l, sh_off, sh_delta);
  const MGLevel& L = m.lv[l];
  RowCtx rc = row_ctx(L.lanes, L.count);
  coarse_row<OP, H, WC>(m, l, source, k, rc, sh_off, sh_delta, prod);
}
// op 8 into level l (from l-1) with G lanes per coarse node over the fine
// support.  The reference adds the active children in child order
// (pressure_multigrid.glslinc:200-224); adding +0 for an inactive child
// instead leaves every partial sum unchanged (the sum starts at +0, so it is
// never -0), which lets every child contribute unconditionally.  With G > 1 the
// lanes' weighted children go to the node's LDS slice (node-major, child t at
// part[r * MG_RESTRICT_STRIDE + t]) and lane 0 adds them in child order, four
// per LDS read, so the sum is bit-identical to one thread walking the support.
#define MG_RESTRICT_MAX 64      // children of a node: a 4x4x4 (trilinear) or 2x2x2 (aggregation) box
#define MG_RESTRICT_STRIDE 68   // slice stride: float4-aligned, successive nodes four banks apart
// With `fuse` the node's first Chebyshev sweep follows from the b just summed
// (cheb_first).
__global__ void __launch_bounds__(256) k_mg_restrict_coop(Ctx s, MGDev m, int l, int phase, int fuse) {
  if (!mg_enabled(s, phase)) return;   // uniform over the launch: no thread reaches a barrier
  extern __shared__ float part[];      // (256 / G) node slices when G > 1
  const MGLevel& C = m.lv[l]; int parent = l - 1; const MGLevel& P = m.lv[parent];
  RowCtx rc = row_ctx(C.rlanes, C.count);
  float* mine = part + ((int)threadIdx.x / rc.G) * MG_RESTRICT_STRIDE;
  int total = 0;
  float sum = 0;
  I3 nf = mdims(m, parent), nl = mdims(m, l);
  if (rc.valid && mactive(m, l, rc.row)) {
    I3 qc = coord(rc.row, nl);
    int lo[3], ext[3] = {1, 1, 1}, nfa[3] = {nf.x, nf.y, nf.z}, nla[3] = {nl.x, nl.y, nl.z}, qca[3] = {qc.x, qc.y, qc.z};
    for (int a = 0; a < 3; a++) lo[a] = qca[a] * P.f[a];
    if (P.aggregate) { ext[0] = P.f[0]; ext[1] = P.f[1]; ext[2] = P.f[2]; }
    else for (int a = 0; a < 3; a++) if (P.f[a] > 1) {
      lo[a]--; ext[a] = 4;
      if ((a == 2 || (a == 0 && m.perx)) && nfa[a] <= 4) { lo[a] = 0; ext[a] = nfa[a]; }
    }
    total = ext[0] * ext[1] * ext[2];
    // the flag and value loads of a child do not wait on its flag (in-range
    // index for an outside child, which adds nothing), so the loads of
    // several children are in flight together
#pragma unroll 4
    for (int t = rc.lane; t < total; t += rc.G) {
      int x = t % ext[0], y = (t / ext[0]) % ext[1], z = t / (ext[0] * ext[1]);
      const int jw = mwrap(m, parent, mk(lo[0] + x, lo[1] + y, lo[2] + z)), j = jw >= 0 ? jw : 0;
      const float flag = m.vec[at(m, parent, 3, j)], v = m.vec[at(m, parent, 6, j)];
      const bool on = jw >= 0 && flag > .5f;  // mactive(m, parent, jw)
      float weight = 1;
      if (!P.aggregate) {
        I3 q = coord(j, nf); int qa[3] = {q.x, q.y, q.z};
        for (int a = 0; a < 3; a++) {
          float pos = lattice_pos(qa[a], P.f[a]); int base = (int)floorf(pos); float tt = fract(pos);
          float w = 0;
          for (int bit = 0; bit < P.f[a]; bit++) {
            int nb = base + bit;
            if (a == 2 || (a == 0 && m.perx)) nb = pmod(nb, nla[a]);
            else if (a == 0) nb = iclamp(nb, 0, nla[a] - 1); else nb = nb > 0 ? nb : 0;
            if (nb == qca[a]) w += P.f[a] == 1 ? 1 : (bit == 1 ? tt : 1 - tt);
          }
          weight *= w;
        }
      }
      const float pv = on ? weight * v : 0.0f;
      if (rc.G == 1) sum += pv;
      else mine[t] = pv;
    }
  }
  if (rc.G > 1) {
    __syncthreads();
    if (rc.valid && rc.lane == 0) {
      const float4* m4 = reinterpret_cast<const float4*>(mine);
      int t = 0;
      for (; t + 4 <= total; t += 4) { const float4 q = m4[t >> 2]; sum += q.x; sum += q.y; sum += q.z; sum += q.w; }
      for (; t < total; t++) sum += mine[t];
    }
  }
  if (!rc.valid || rc.lane != 0) return;
  m.vec[at(m, l, 0, rc.row)] = sum; m.vec[at(m, l, 1, rc.row)] = 0;
  if (fuse) cheb_first(&m.vec[at(m, l, 7, rc.row)], &m.vec[at(m, l, 2, rc.row)], m.vec[at(m, l, 4, rc.row)], sum);
  else m.vec[at(m, l, 2, rc.row)] = 0;
}
static size_t restrict_lds(const MGDev& m, int l) {
  const int G = m.lv[l].rlanes;
  return G > 1 ? sizeof(float) * (256 / G) * MG_RESTRICT_STRIDE : 0;
}

// terminal smoother when coarsening stops above 64 nodes
// (water/pressure_multigrid.gd:144-149): op 14 then op 6 sweeps
__global__ void __launch_bounds__(256) k_mg_jacobi(Ctx s, MGDev m, int l, int source, int first, int phase) {
  if (!mg_enabled(s, phase)) return;
  int64_t i = GID(); if (i >= m.lv[l].count) return;
  if (first) { m.vec[at(m, l, 2, i)] = m.vec[at(m, l, 4, i)] * m.vec[at(m, l, 0, i)]; return; }
  m.vec[at(m, l, 3 - source, i)] = m.vec[at(m, l, source, i)] + m.vec[at(m, l, 4, i)] * (m.vec[at(m, l, 0, i)] - mapply(m, l, source, (int)i));
}
// op 7
template <bool H>
__global__ void __launch_bounds__(256) k_mg_residual(Ctx s, MGDev m, int l, int field, int phase) {
  if (!mg_enabled(s, phase)) return;
  int64_t i = GID(); if (i >= m.lv[l].count) return;
  m.vec[at(m, l, 6, i)] = m.vec[at(m, l, 0, i)] - mapply0<H>(m, field, (int)i);
}
// op 9: x_l += P x_{l+1}
__device__ __forceinline__ void prolong_row(const MGDev& m, int l, int ff, int cf, int64_t i) {
  float add = 0;
  const MGLevel& L = m.lv[l];
  if (L.aggregate) {
    if (mactive(m, l, (int)i)) { int ids[8]; float ws[8]; int size = mtransfers(m, l, (int)i, ids, ws); for (int k = 0; k < size; k++) add += ws[k] * m.vec[at(m, l + 1, cf, ids[k])]; }
  } else {
    // mtransfers' parents (at most 2 x 2 x 2: factors are 1 or 2) with their
    // flags and values loaded together, then added in its order (x fastest)
    // under its conditions
    const bool on = m.vec[at(m, l, 3, i)] > .5f;  // mactive(m, l, i)
    const I3 n = mdims(m, l + 1), q = coord(i, mdims(m, l));
    int base[3]; float t[3]; const int qq[3] = {q.x, q.y, q.z};
    for (int k = 0; k < 3; k++) { float pos = lattice_pos(qq[k], L.f[k]); base[k] = (int)floorf(pos); t[k] = fract(pos); }
    float wc[8], fc[8], vc[8]; bool use[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int bit[3] = {c & 1, (c >> 1) & 1, c >> 2};
      I3 nb = mk(base[0] + bit[0], base[1] + bit[1], base[2] + bit[2]); float ww = 1;
      for (int axis = 0; axis < 3; axis++) if (L.f[axis] > 1) ww *= bit[axis] == 1 ? t[axis] : 1 - t[axis];
      use[c] = bit[0] < L.f[0] && bit[1] < L.f[1] && bit[2] < L.f[2] && nb.y < n.y;
      if (nb.y < 0) nb.y = 0;
      nb.x = m.perx ? pmod(nb.x, n.x) : iclamp(nb.x, 0, n.x - 1); nb.z = pmod(nb.z, n.z);
      const int j = use[c] ? (int)index3(nb, n) : 0;
      wc[c] = ww; fc[c] = m.vec[at(m, l + 1, 3, j)]; vc[c] = m.vec[at(m, l + 1, cf, j)];
    }
    if (on) {
#pragma unroll
      for (int c = 0; c < 8; c++) if (use[c] && fc[c] > .5f && wc[c] != 0) add += wc[c] * vc[c];
    }
  }
  m.vec[at(m, l, ff, i)] += add;
}
__global__ void __launch_bounds__(256) k_mg_prolong(Ctx s, MGDev m, int l, int ff, int cf, int phase) {
  if (!mg_enabled(s, phase)) return;
  int64_t i = GID(); if (i >= m.lv[l].count) return;
  prolong_row(m, l, ff, cf, i);
}
// pressure_mg_coarse.glslinc:50-53 (op 12): z = L^-T (L^-1 b), one group
__global__ void __launch_bounds__(256) k_mg_coarse_solve(Ctx s, MGDev m, int l, int phase) {
  if (!mg_enabled(s, phase)) return;
  __shared__ float rhs[64];
  int lid = threadIdx.x, N = m.lv[l].count; const float* Li = m.linv;
  if (lid < N) { float d = 0; for (int j = 0; j <= lid; j++) d += Li[lid * N + j] * m.vec[at(m, l, 0, j)]; rhs[lid] = d; }
  __syncthreads();
  if (lid < N) { float d = 0; for (int j = lid; j < N; j++) d += Li[j * N + lid] * rhs[j]; m.vec[at(m, l, 1, lid)] = d; }
}
// pressure_multigrid.glslinc:230-234 (op 10) fused with the PCG finaliser:
// phase 0 -> pressure_iterations.glslinc:14-19 (op 4), phase 1 -> :31 (op 9)
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_mg_out(Ctx s, MGDev m, int field, int phase, int) {
  if (!mg_enabled(s, phase)) return;
  if (!FINAL) {
  int64_t i = GID(); int64_t N = s.N; float4 value = make_float4(0, 0, 0, 0);
  if (i < N) {
    float r = s.work[i], z = m.vec[at(m, 0, field, i)], b = s.work[5 * N + i];
    s.work[N + i] = z; if (phase == 0) s.work[2 * N + i] = z;
    value = make_float4(r * z, r * r, b * b, fabsf(r));
  }
  value = group_reduce<RED_PRESSURE>(value);
  if (threadIdx.x == 0) s.partial[blockIdx.x] = value;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float4 v = global_reduce<true>(s.partial, G);
  if (threadIdx.x != 0) return;
  float* c = s.c;
  if (phase == 0) {
    c[9] = v.x; c[10] = v.z; c[13] = sqrtf(v.y / gmax(v.z, 1e-30f));
    if (residual_converged(s, v, v.z)) c[6] = 1;
    if (c[58] == 0) c[59] = v.z;
    c[57] = 0;
  } else {
    c[12] = v.x / gmax(c[9], 1e-30f); c[9] = v.x; c[8]++; c[18]++;
    c[13] = sqrtf(v.y / gmax(c[10], 1e-30f));
    if (isnan(c[13]) || isinf(c[13])) c[7] = 2;
    else if (residual_converged(s, v, c[10])) c[6] = 1;
    else if (c[8] > 4096) c[7] = 3;
  }
}
// op 3 when the V-cycle is predicated off (a failure code is set): the
// reference's op 4 then reduces op 3's own partials.
template <bool FINAL>
__global__ void __launch_bounds__(WV_RED) k_pcg_init_partials(Ctx s, int) {
  if (!pressure_setup(s)) return;
  if (mg_enabled(s, 0)) return;
  if (!FINAL) {
  int64_t i = GID(); int64_t N = s.N; float4 v = make_float4(0, 0, 0, 0);
  if (i < N) { float r = s.work[i], z = s.work[N + i], bv = s.work[5 * N + i]; v = make_float4(r * z, r * r, bv * bv, fabsf(r)); }
  v = group_reduce<RED_PRESSURE>(v);
  if (threadIdx.x == 0) s.partial[blockIdx.x] = v;
  return;
  }
  const int64_t G = (s.N + WV_RED - 1) / WV_RED;
  float4 g = global_reduce<true>(s.partial, G);
  if (threadIdx.x != 0) return;
  float* c = s.c;
  c[9] = g.x; c[10] = g.z; c[13] = sqrtf(g.y / gmax(g.z, 1e-30f));
  if (residual_converged(s, g, g.z)) c[6] = 1;
  if (c[58] == 0) c[59] = g.z;
  c[57] = 0;
}

// ---------------------------------------------------- hierarchy audit -------
// pressure_hierarchy_state.glsl:60-68
__device__ __forceinline__ unsigned signature(const Ctx& s, int64_t i) {
  if (s.geo[i].y < .5f) return 0;
  I3 n = dims(s), q = coord(i, n); unsigned result = 0;
  for (int f6 = 0; f6 < 6; f6++) {
    int axis = f6 >> 1, side = f6 & 1; I3 nb = q; addc(nb, axis, 2 * side - 1);
    nb.z = pmod(nb.z, n.z); if (s.perx > .5f) nb.x = pmod(nb.x, n.x);
    bool inside = nb.x >= 0 && nb.x < n.x && nb.y >= 0 && nb.y < n.y;
    bool w = inside && s.geo[index3(nb, n)].y > .5f;
    unsigned kind = s.work[(int64_t)(6 + f6) * s.N + i] > 0 ? (w ? 1u : 2u) : 0u;
    result |= kind << (2 * f6);
  }
  return result;
}
constexpr int AUDIT_STRIPES = 64;  // stripes of 16 counters (a cache line each) in MGDev::audit_part
__device__ __forceinline__ bool audit_on(const Ctx& s) { return !(s.c[5] < .5f || s.c[7] > .5f); }
__global__ void __launch_bounds__(256) k_audit0(Ctx s, MGDev m) {
  if (!audit_on(s)) return;
  for (int k = threadIdx.x; k < 16 * AUDIT_STRIPES; k += blockDim.x) m.audit_part[k] = 0;
  if (threadIdx.x != 0) return;
  unsigned* st = m.state;
  for (int k = 1; k <= 7; k++) st[k] = 0; st[14] = st[0]; st[15] = 0; st[16] = 0; for (int k = 20; k < 32; k++) st[k] = 0;
}
// Counts of the audit.  Each thread reads everything its cell needs up front
// (the neighbours' kinds at a clamped index when outside, masked afterwards),
// the wave and the block sum the counts (integers; the ratio as its float bits
// under an unsigned maximum, as the atomic maximum of the reference takes it)
// and one thread adds the block's totals into one of AUDIT_STRIPES padded
// stripes, which k_audit2 sums into the header: the totals are those of the
// per-cell atomics, without their contention on one cache line.
__global__ void __launch_bounds__(256) k_audit1(Ctx s, MGDev m) {
  if (!audit_on(s)) return;
  __shared__ unsigned part[4][9];
  // a_x, a_y, a_z, a_w, b_x, b_y, ratio bits (maximum), d_x, d_y
  unsigned acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t n = s.N, i = GID();
  if (i < n) {
    const float* cached = m.snapshot; const float* prev = m.previous;
    const I3 dn = dims(s), q = coord(i, dn);
    float value[7], was[7], pw[7], ngy[6];
#pragma unroll
    for (int k = 0; k < 7; k++) { value[k] = s.work[(6 + k) * n + i]; was[k] = cached[k * n + i]; pw[k] = prev[k * n + i]; }
    const float kind = s.geo[i].y, ck = cached[7 * n + i], pk = prev[7 * n + i];
    const unsigned old = (unsigned)cached[8 * n + i], pv = (unsigned)prev[8 * n + i];
    bool inside[6];
#pragma unroll
    for (int f6 = 0; f6 < 6; f6++) {
      // pressure_hierarchy_state.glsl:60-68 (the signature's neighbour of face f6)
      const int axis = f6 >> 1, side = f6 & 1; I3 nb = q; addc(nb, axis, 2 * side - 1);
      nb.z = pmod(nb.z, dn.z); if (s.perx > .5f) nb.x = pmod(nb.x, dn.x);
      inside[f6] = nb.x >= 0 && nb.x < dn.x && nb.y >= 0 && nb.y < dn.y;
      ngy[f6] = s.geo[inside[f6] ? index3(nb, dn) : i].y;
    }
    unsigned sig = 0;
    if (!(kind < .5f)) {
#pragma unroll
      for (int f6 = 0; f6 < 6; f6++) {
        const bool w = inside[f6] && ngy[f6] > .5f;
        const unsigned kd = value[f6] > 0 ? (w ? 1u : 2u) : 0u;
        sig |= kd << (2 * f6);
      }
    }
    acc[0] = kind != ck; acc[3] = kind != pk; acc[8] = kind > .5f;
#pragma unroll
    for (int face = 0; face < 6; face++) { acc[1] += ((sig >> (2 * face)) & 3) != ((old >> (2 * face)) & 3); acc[4] += ((sig >> (2 * face)) & 3) != ((pv >> (2 * face)) & 3); }
    float ratio = 0; unsigned d_x = 0;
#pragma unroll
    for (int k = 0; k < 7; k++) {
      acc[2] += __float_as_uint(value[k]) != __float_as_uint(was[k]); acc[5] += __float_as_uint(value[k]) != __float_as_uint(pw[k]);
      if (was[k] > 0) ratio = gmax(ratio, fabsf(value[k] - was[k]) / was[k]);
      if (isnan(value[k]) || isinf(value[k]) || value[k] < 0) d_x = 1;
    }
    acc[6] = __float_as_uint(ratio); acc[7] = d_x;
  }
#pragma unroll
  for (int k = 0; k < 9; k++) {
    unsigned v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      unsigned o = (unsigned)__shfl_xor((int)v, off, 64);
      v = k == 6 ? (v > o ? v : o) : v + o;
    }
    acc[k] = v;
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 9; k++) part[wave][k] = acc[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned* stripe = m.audit_part + 16 * (blockIdx.x % AUDIT_STRIPES);
    for (int k = 0; k < 9; k++) {
      unsigned v = part[0][k];
      for (int w = 1; w < 4; w++) v = k == 6 ? (v > part[w][k] ? v : part[w][k]) : v + part[w][k];
      if (k == 6) atomicMax(&stripe[k], v); else if (v) atomicAdd(&stripe[k], v);
    }
  }
}
__global__ void __launch_bounds__(256) k_audit2(Ctx s, MGDev m) {
  if (!audit_on(s)) return;
  unsigned* st = m.state;
  {  // the stripes of k_audit1 into the header (one stripe per lane)
    const int lane = threadIdx.x;
    unsigned v[9];
    for (int k = 0; k < 9; k++) v[k] = lane < AUDIT_STRIPES ? m.audit_part[16 * lane + k] : 0u;
    for (int k = 0; k < 9; k++)
      for (int off = 32; off > 0; off >>= 1) {
        unsigned o = (unsigned)__shfl_xor((int)v[k], off, 64);
        v[k] = k == 6 ? (v[k] > o ? v[k] : o) : v[k] + o;
      }
    if (lane != 0) return;
    for (int k = 0; k < 4; k++) st[1 + k] += v[k];
    st[5] += v[4]; st[6] += v[5]; st[7] = st[7] > v[6] ? st[7] : v[6]; st[15] += v[7]; st[16] += v[8];
  }
  if (st[15] > 0) { s.c[7] = 34; return; }
  int mode = m.mode;
  bool incompatible = st[0] == 0 || st[1] > 0 || st[2] > 0;
  bool structural = mode == 0 || incompatible;
  bool numerical = mode == 0 || incompatible || (st[3] > 0 && (mode == 1 || st[10] + 1 >= (unsigned)m.lag));
  st[8] = numerical; st[9] = structural; st[10] = numerical || st[3] == 0 ? 0 : st[10] + 1; st[11] += st[8]; st[12] += st[9]; st[13]++; st[0] = 1;
  if (st[13] <= (unsigned)m.capacity) {
    int64_t row = 64 + 32 * ((int64_t)st[13] - 1); st[row] = st[13]; st[row + 1] = __float_as_uint(s.c[1]);
    for (int k = 1; k <= 7; k++) st[row + 1 + k] = st[k];
    st[row + 9] = st[8]; st[row + 10] = st[9]; st[row + 11] = st[10]; st[row + 12] = st[16]; st[row + 13] = st[11];
  }
}
__global__ void __launch_bounds__(256) k_audit3(Ctx s, MGDev m) {
  if (!audit_on(s)) return;
  int64_t i = GID(), n = s.N; if (i >= n) return;
  bool numeric = m.state[8] != 0;
  for (int k = 0; k < 7; k++) { float value = s.work[(6 + k) * n + i]; m.previous[k * n + i] = value; if (numeric) m.snapshot[k * n + i] = value; }
  float kind = s.geo[i].y, sig = (float)signature(s, i);
  m.previous[7 * n + i] = kind; m.previous[8 * n + i] = sig;
  if (numeric) { m.snapshot[7 * n + i] = kind; m.snapshot[8 * n + i] = sig; }
}
__global__ void __launch_bounds__(256) k_audit4(Ctx s, MGDev m) {
  if (!audit_on(s)) return;
  unsigned* st = m.state;
  if (st[13] <= (unsigned)m.capacity) { int64_t row = 64 + 32 * ((int64_t)st[13] - 1); for (int l = 0; l < 12; l++) st[row + 16 + l] = st[20 + l]; }
}

// ------------------------------------------------------------ MG build ------
__device__ __forceinline__ bool build_on(const Ctx& s, const MGDev& m, bool structural) {
  if (s.c[7] > .5f || s.c[5] < .5f) return false;
  return structural ? m.state[9] != 0 : m.state[8] != 0;
}
// pressure_multigrid.glslinc:139-150 (op 1)
__global__ void __launch_bounds__(256) k_mg_classify(Ctx s, MGDev m, int l) {
  if (!build_on(s, m, true)) return;
  int64_t i = GID(); if (i >= m.lv[l].count) return;
  float value = 0;
  if (l == 0) value = s.geo[i].y > .5f ? 1 : 0;
  else {
    const MGLevel& P = m.lv[l - 1]; I3 base = coord(i, mdims(m, l));
    base.x *= P.f[0]; base.y *= P.f[1]; base.z *= P.f[2];
    for (int z = 0; z < P.f[2]; z++) for (int y = 0; y < P.f[1]; y++) for (int x = 0; x < P.f[0]; x++)
      if (mactive(m, l - 1, mwrap(m, l - 1, mk(base.x + x, base.y + y, base.z + z)))) value = 1;
  }
  // changed flags counted per wave, one atomic per wave (an integer count)
  const bool changed = m.vec[at(m, l, 3, i)] != value;
  const unsigned long long ballot = __ballot(changed);
  if (changed && __lane_id() == __ffsll((long long)ballot) - 1) atomicAdd(&m.state[20 + l], (unsigned)__popcll(ballot));
  m.vec[at(m, l, 3, i)] = value;
}
// pressure_multigrid.glslinc:183-189 (op 4), level 0
__global__ void __launch_bounds__(256) k_mg_l1_fine(Ctx s, MGDev m) {
  if (!build_on(s, m, false)) return;
  int64_t i = GID(); if (i >= m.lv[0].count) return;
  I3 q = coord(i, mdims(m, 0)); float sum = snap(m, 12, i);
  // the six neighbours' flags and coefficients loaded together (in-range
  // index for an outside neighbour), then added in face order
  float flag[6], cf[6]; bool in[6];
#pragma unroll
  for (int f6 = 0; f6 < 6; f6++) {
    int axis = f6 >> 1, side = f6 & 1; I3 nb = q; addc(nb, axis, 2 * side - 1);
    const int j = mwrap(m, 0, nb); in[f6] = j >= 0;
    flag[f6] = m.vec[at(m, 0, 3, j >= 0 ? j : 0)]; cf[f6] = snap(m, 6 + f6, i);
  }
#pragma unroll
  for (int f6 = 0; f6 < 6; f6++) if (in[f6] && flag[f6] > .5f) sum += cf[f6];  // mactive(m, 0, j)
  if (!(sum > 0) || isnan(sum) || isinf(sum)) { s.c[7] = 32; return; }
  m.vec[at(m, 0, 4, i)] = 1 / sum;
}
// op 4 at a coarse level: l1 row sums of the (symmetric) Galerkin rows
__global__ void __launch_bounds__(256) k_mg_l1_coarse(Ctx s, MGDev m, int l) {
  if (!build_on(s, m, false)) return;
  int64_t i = GID(); int N = m.lv[l].count; if (i >= N) return;
  const float* row = m.rows + m.lv[l].row_off + i; float sum = 0;
  // the row's entries in order, eight loads in flight ahead of their additions
  const int W = m.lv[l].width;
  int k = 0;
  for (; k + 8 <= W; k += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = row[(int64_t)(k + u) * N];
#pragma unroll
    for (int u = 0; u < 8; u++) sum += fabsf(v[u]);
  }
  for (; k < W; k++) sum += fabsf(row[(int64_t)k * N]);
  if (!(sum > 0) || isnan(sum) || isinf(sum)) { s.c[7] = 32; return; }
  m.vec[at(m, l, 4, i)] = 1 / sum;
}

// ---- Galerkin P^T A P in two gather stages ----------------------------------
// Stage 1, one thread per fine node i of level l: the row (A P)_i, i.e.
//   T_i[J] = sum_j A[i, j] P[j, J]
// kept as a small dense box over the coarse nodes J the row can reach, indexed
// by the parents' unclamped coordinates relative to an anchor (so the box is a
// fixed shape per level).  Stage 2, G lanes per coarse node I of level l+1:
//   C[I, J] = sum_i P[i, I] T_i[J]
// over the fine nodes i whose transfers reach I.  Both stages run their sums in
// a fixed order: the coarse operator is deterministic.
__device__ __forceinline__ int floordiv(int a, int b) {
  if (b == 1) return a;
  if (b == 2) return a >> 1;  // arithmetic shift: the floor of a / 2
  int q = a / b; return (a % b != 0 && (a < 0) != (b < 0)) ? q - 1 : q;
}
// first (unclamped) parent coordinate of fine coordinate j along axis a
__device__ __forceinline__ int parent_base(const MGLevel& F, int a, int j) {
  if (F.aggregate) return floordiv(j, F.f[a]);
  if (F.f[a] == 1) return j;
  return (int)floorf(lattice_pos(j, F.f[a]));
}
// coarse node of an unclamped parent coordinate (pressure_multigrid.glslinc:115-117),
// -1 if dropped (beyond the top)
__device__ __forceinline__ int parent_node(const MGDev& m, int lc, const int P[3]) {
  I3 n = mdims(m, lc);
  int y = P[1]; if (y >= n.y) return -1; if (y < 0) y = 0;
  int x = m.perx ? pmod(P[0], n.x) : iclamp(P[0], 0, n.x - 1);
  int z = pmod(P[2], n.z);
  return (int)index3(mk(x, y, z), n);
}
// G lanes per fine row (a power of two dividing 128): lane g takes entries
// g, g + G, ... into its own slice of the accumulator, and lane 0 folds the
// slices in lane order into (A P)_i.
__global__ void __launch_bounds__(128) k_mg_galerkin_ap(Ctx s, MGDev m, int l, int G) {
  if (!build_on(s, m, false)) return;
  extern __shared__ float acc[];  // box entries x 128 lanes, sized per level
  const MGLevel& F = m.lv[l]; int lc = l + 1;
  const int64_t gid = GID(); const int tid = threadIdx.x, lg = __ffs(G) - 1, lane = tid & (G - 1);
  const int64_t i = gid >> lg;
  int E0 = F.box[0], E1 = F.box[1], E = F.box[0] * F.box[1] * F.box[2];
  for (int b = 0; b < E; b++) acc[b * 128 + tid] = 0;
  const bool live = i < F.count && mactive(m, l, (int)i);
  if (live) {
  I3 q = coord(i, mdims(m, l)); int qa[3] = {q.x, q.y, q.z};
  int anchor[3]; for (int a = 0; a < 3; a++) anchor[a] = parent_base(F, a, qa[a] - F.margin[a]);
  // bit b: the coarse node at box entry b = bx + E0 (by + E1 bz) exists and
  // is active.  Every parent the row's entries reach lies in the box, so the
  // flags are loaded once per row, eight at a time, not once per (entry,
  // parent) pair.
  uint64_t live_parents = 0;
  {
    int bx = 0, by = 0, bz = 0;
    for (int b0 = 0; b0 < E; b0 += 8) {
      int J[8]; float fl[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int P[3] = {anchor[0] + bx, anchor[1] + by, anchor[2] + bz};
        J[u] = b0 + u < E ? parent_node(m, lc, P) : -1;
        fl[u] = m.vec[at(m, lc, 3, J[u] >= 0 ? J[u] : 0)];
        if (++bx == E0) { bx = 0; if (++by == E1) { by = 0; ++bz; } }
      }
#pragma unroll
      for (int u = 0; u < 8; u++) if (J[u] >= 0 && fl[u] > .5f) live_parents |= 1ull << (b0 + u);
    }
  }
  const int width = F.width;
  // the row's entries k = lane, lane + G, ... eight at a time: their offsets,
  // coefficients and neighbour flags loaded together, then accumulated in k
  // order
  for (int k0 = lane; k0 < width; k0 += 8 * G) {
    int o8[8][3], j8[8]; float a8[8], f8[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int k = k0 + u * G;
      int off[3] = {0, 0, 0}; float aij = 0; int j = -1;
      if (k < width) {
        if (l == 0) {
          if (k == 0) { aij = snap(m, 12, i); j = (int)i; }
          else { int f6 = k - 1, axis = f6 >> 1, side = f6 & 1; off[axis] = 2 * side - 1; aij = -snap(m, 6 + f6, i);
                 j = mwrap(m, 0, mk(q.x + off[0], q.y + off[1], q.z + off[2])); }
        } else {
          I3 o = moffset(m, l, k); off[0] = o.x; off[1] = o.y; off[2] = o.z;
          j = mwrap(m, l, mk(q.x + o.x, q.y + o.y, q.z + o.z));
          const float a_row = m.rows[F.row_off + (int64_t)k * F.count + i];  // loaded before j is known
          aij = j < 0 ? 0 : a_row;
        }
      }
      // the neighbour's flag at an in-range index (mactive(m, l, j) for j >= 0)
      f8[u] = m.vec[at(m, l, 3, j >= 0 ? j : 0)];
      o8[u][0] = off[0]; o8[u][1] = off[1]; o8[u][2] = off[2]; j8[u] = j; a8[u] = aij;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const float aij = a8[u];
      if (j8[u] < 0 || aij == 0 || !(f8[u] > .5f)) continue;
      int ju[3] = {qa[0] + o8[u][0], qa[1] + o8[u][1], qa[2] + o8[u][2]};  // unwrapped neighbour
      int base[3]; float t[3];
      for (int a = 0; a < 3; a++) {
        if (F.aggregate || F.f[a] == 1) { base[a] = parent_base(F, a, ju[a]); t[a] = 0; }
        else { float pos = lattice_pos(ju[a], F.f[a]); base[a] = (int)floorf(pos); t[a] = fract(pos); }
      }
      // at most 2 x 2 x 2 parents (coarsening factors are 1 or 2), accumulated
      // in x-fastest order; a parent's flag is its bit in live_parents
      const int nb0 = F.aggregate ? 1 : F.f[0], nb1 = F.aggregate ? 1 : F.f[1], nb2 = F.aggregate ? 1 : F.f[2];
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int x = c & 1, y = (c >> 1) & 1, z = c >> 2;
        if (!(x < nb0 && y < nb1 && z < nb2)) continue;
        int bit[3] = {x, y, z}; float w = 1;
#pragma unroll
        for (int a = 0; a < 3; a++) if (!F.aggregate && F.f[a] > 1) w *= bit[a] == 1 ? t[a] : 1 - t[a];
        const int bc = (base[0] + x - anchor[0]) + E0 * ((base[1] + y - anchor[1]) + E1 * (base[2] + z - anchor[2]));
        if (w == 0 || !((live_parents >> bc) & 1)) continue;
        acc[bc * 128 + tid] += aij * w;
      }
    }
  }
  }
  __syncthreads();
  if (!live) return;
  // entry b of the row folds its G lane slices in lane order; the row's lanes
  // share the entries (lane g: quads g, g + G, ...)
  const int t0 = tid - lane;
  auto box = [&](int b) { if (b >= E) return 0.f; float v = acc[b * 128 + t0]; for (int g = 1; g < G; g++) v += acc[b * 128 + t0 + g]; return v; };
  // (A P)_i is stored row-major, padded to a multiple of 4 floats, so that the
  // second stage reads a fine row's box with 16-byte loads
  const int Ep = (E + 3) & ~3;
  float4* T = (float4*)(m.ap + i * Ep);
  for (int b = 4 * lane; b < Ep; b += 4 * G) T[b >> 2] = make_float4(box(b), box(b + 1), box(b + 2), box(b + 3));
}
#define GAL_THREADS 64
__global__ void __launch_bounds__(GAL_THREADS) k_mg_galerkin_ptap(Ctx s, MGDev m, int l) {
  if (!build_on(s, m, false)) return;
  extern __shared__ float acc[];  // width(l+1) * GAL_THREADS, one slice per lane
  int lc = l + 1; const MGLevel& C = m.lv[lc]; const MGLevel& F = m.lv[l];
  // lanes per coarse node over its fine support: the restriction's split, at
  // most 8 (lane 0 folds the G accumulator slices)
  const int G = C.rlanes > 8 ? 8 : C.rlanes;
  int64_t gid = GID(); int tid = threadIdx.x;
  int64_t I = gid >> (__ffs(G) - 1); int lane = (int)(gid & (G - 1));
  int W = C.width;
  for (int k = 0; k < W; k++) acc[k * GAL_THREADS + tid] = 0;
  // the level's slot table (delta -> stencil entry) in LDS: looked up once per
  // box entry, right before the accumulator update that needs it
  __shared__ signed char slots[128];
  for (int k = tid; k < 125; k += GAL_THREADS) slots[k] = m.slots[125 * lc + k];
  __syncthreads();
  bool live = I < C.count && mactive(m, lc, (int)I);
  int E0 = F.box[0], E1 = F.box[1], E2 = F.box[2], E = E0 * E1 * E2;
  if (live) {
    I3 qc = coord(I, mdims(m, lc)); I3 nf = mdims(m, l);
    int lo[3], hi[3]; int qca[3] = {qc.x, qc.y, qc.z}; int nfa[3] = {nf.x, nf.y, nf.z};
    for (int a = 0; a < 3; a++) {
      if (F.aggregate) { lo[a] = qca[a] * F.f[a]; hi[a] = lo[a] + F.f[a] - 1; if (hi[a] > nfa[a] - 1) hi[a] = nfa[a] - 1; }
      else if (F.f[a] == 1) { lo[a] = qca[a]; hi[a] = qca[a]; }
      else { lo[a] = 2 * qca[a] - 1; hi[a] = 2 * qca[a] + 2;
        bool periodic = a == 2 || (a == 0 && m.perx);
        if (!periodic) { if (lo[a] < 0) lo[a] = 0; if (hi[a] > nfa[a] - 1) hi[a] = nfa[a] - 1; }
      }
    }
    int sx = hi[0] - lo[0] + 1, sy = hi[1] - lo[1] + 1, sz = hi[2] - lo[2] + 1, total = sx * sy * sz;
    for (int c = lane; c < total; c += G) {
      int x = lo[0] + c % sx, y = lo[1] + (c / sx) % sy, z = lo[2] + c / (sx * sy);
      int i = mwrap(m, l, mk(x, y, z)); if (i < 0 || !mactive(m, l, i)) continue;
      int ids[8]; float ws[8]; int ni = mtransfers(m, l, i, ids, ws);
      float wI = 0; for (int a = 0; a < ni; a++) if (ids[a] == (int)I) wI += ws[a];
      if (wI == 0) continue;
      I3 q = coord(i, nf); int qa[3] = {q.x, q.y, q.z};
      int anchor[3]; for (int a = 0; a < 3; a++) anchor[a] = parent_base(F, a, qa[a] - F.margin[a]);
      const float4* T = (const float4*)(m.ap + (int64_t)i * ((E + 3) & ~3));
      // box entry b = bx + E0 (by + E1 bz); its coarse node's offset from I is
      // taken from the parent coordinates directly (mdelta of the encoded node)
      I3 nc = mdims(m, lc);
      // box entry b = bx + E0 (by + E1 bz), x fastest, walked a quad at a
      // time with the next quad's load in flight, so that the load latency
      // overlaps the accumulator updates
      int bx = 0, by = 0, bz = 0;
      auto entry = [&](float v) {
        const int ex = bx, ey = by, ez = bz;
        if (++bx == E0) { bx = 0; if (++by == E1) { by = 0; ++bz; } }
        if (v == 0) return;
        int y = anchor[1] + ey;
        if (y >= nc.y) { s.c[7] = 31; return; }   // parent_node's dropped node: no slot
        if (y < 0) y = 0;
        int x = m.perx ? pmod(anchor[0] + ex, nc.x) : iclamp(anchor[0] + ex, 0, nc.x - 1);
        int z = pmod(anchor[2] + ez, nc.z);
        I3 d = mk(x - qc.x, y - qc.y, z - qc.z);
        if (m.perx) d.x = pmod(d.x + nc.x / 2, nc.x) - nc.x / 2;
        d.z = pmod(d.z + nc.z / 2, nc.z) - nc.z / 2;
        int sl = in_box2(d) ? slots[slot125(d)] : -1;
        if (sl < 0) { s.c[7] = 31; return; }
        acc[sl * GAL_THREADS + tid] += wI * v;
      };
      const int q4 = (E + 3) >> 2;
      float4 cur = T[0];
      for (int q = 0; q < q4; q++) {
        const float4 nxt = T[q + 1 < q4 ? q + 1 : q];
        entry(cur.x);
        if (4 * q + 1 < E) entry(cur.y);
        if (4 * q + 2 < E) entry(cur.z);
        if (4 * q + 3 < E) entry(cur.w);
        cur = nxt;
      }
    }
  }
  __syncthreads();
  if (I >= C.count) return;
  // entry k folds the node's G lane slices in lane order; the node's lanes
  // share the entries (lane g: entries g, g + G, ...)
  const int t0 = tid - lane;
  float* row = m.rows + C.row_off + I;
  for (int k = lane; k < W; k += G) {
    float v;
    if (live) { v = 0; for (int g = 0; g < G; g++) v += acc[k * GAL_THREADS + t0 + g]; }
    else { I3 off = moffset(m, lc, k); v = (off.x == 0 && off.y == 0 && off.z == 0) ? 1.0f : 0.0f; }
    row[(int64_t)k * C.count] = v;
  }
}
// Stage 2 as a gather per coarse entry: block b is coarse node I, lane k owns
// the entry at stencil offset k and adds, over the fine support i of I (in
// support order), w_iI times the box entries of (A P)_i whose coarse node is
// J = I + offset k: one box position in the interior, an interval where the
// transfer clamps (x without periodicity, y below 0), a period apart along a
// periodic axis.  The entry counts only if offset k is J's minimal-image
// offset (the slot the box entry maps to); other aliases of J stay 0 and are
// zeroed by the symmetrisation.  No accumulator in LDS, a fixed order: the
// operator is deterministic.
#define GATHER_SUPPORT 64
__global__ void __launch_bounds__(128) k_mg_galerkin_gather(Ctx s, MGDev m, int l) {
  if (!build_on(s, m, false)) return;
  const int lc = l + 1; const MGLevel& C = m.lv[lc]; const MGLevel& F = m.lv[l];
  const int I = (int)blockIdx.x, k = threadIdx.x, W = C.width;
  __shared__ float wsup[GATHER_SUPPORT];
  __shared__ int isup[GATHER_SUPPORT], asup[GATHER_SUPPORT][3];
  float* row = m.rows + C.row_off + I;
  if (!mactive(m, lc, I)) {  // inactive rows are the identity
    if (k < W) { I3 off = moffset(m, lc, k); row[(int64_t)k * C.count] = (off.x == 0 && off.y == 0 && off.z == 0) ? 1.0f : 0.0f; }
    return;
  }
  const I3 qc = coord(I, mdims(m, lc)), nf = mdims(m, l), nc = mdims(m, lc);
  int lo[3], hi[3]; const int qca[3] = {qc.x, qc.y, qc.z}, nfa[3] = {nf.x, nf.y, nf.z};
  for (int a = 0; a < 3; a++) {
    if (F.aggregate) { lo[a] = qca[a] * F.f[a]; hi[a] = lo[a] + F.f[a] - 1; if (hi[a] > nfa[a] - 1) hi[a] = nfa[a] - 1; }
    else if (F.f[a] == 1) { lo[a] = qca[a]; hi[a] = qca[a]; }
    else { lo[a] = 2 * qca[a] - 1; hi[a] = 2 * qca[a] + 2;
      bool periodic = a == 2 || (a == 0 && m.perx);
      if (!periodic) { if (lo[a] < 0) lo[a] = 0; if (hi[a] > nfa[a] - 1) hi[a] = nfa[a] - 1; }
    }
  }
  const int sx = hi[0] - lo[0] + 1, sy = hi[1] - lo[1] + 1, total = sx * sy * (hi[2] - lo[2] + 1);
  if (k < total) {
    int x = lo[0] + k % sx, y = lo[1] + (k / sx) % sy, z = lo[2] + k / (sx * sy);
    int i = mwrap(m, l, mk(x, y, z)); float wI = 0;
    if (i >= 0 && mactive(m, l, i)) {
      int ids[8]; float ws[8]; int ni = mtransfers(m, l, i, ids, ws);
      for (int a = 0; a < ni; a++) if (ids[a] == I) wI += ws[a];
      I3 q = coord(i, nf); int qa[3] = {q.x, q.y, q.z};
      for (int a = 0; a < 3; a++) asup[k][a] = parent_base(F, a, qa[a] - F.margin[a]);
    }
    wsup[k] = wI; isup[k] = i;
  }
  __syncthreads();
  // the support's (A P) boxes, read once with 16-byte loads (a box is at most
  // 64 floats, contiguous and padded to 4); the entries are summed from LDS.
  // The launch sizes sbox to GATHER_SUPPORT boxes of this level (not of the
  // largest box), which sets how many blocks share a CU.
  const int Ep = (F.box[0] * F.box[1] * F.box[2] + 3) & ~3, q4 = Ep >> 2;
  extern __shared__ float4 sbox[];
  for (int e = k; e < total * q4; e += (int)blockDim.x) {
    const int c = e / q4, r = e - c * q4;
    sbox[c * q4 + r] = wsup[c] != 0 ? reinterpret_cast<const float4*>(m.ap + (int64_t)isup[c] * Ep)[r] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  __syncthreads();
  if (k >= W) return;
  const I3 off = moffset(m, lc, k);
  int J[3] = {qc.x + off.x, qc.y + off.y, qc.z + off.z};
  bool inside = J[1] >= 0 && J[1] < nc.y;
  if (m.perx) J[0] = pmod(J[0], nc.x); else inside = inside && J[0] >= 0 && J[0] < nc.x;
  J[2] = pmod(J[2], nc.z);
  float acc = 0;
  if (inside) {
    I3 d = mk(J[0] - qc.x, J[1] - qc.y, J[2] - qc.z);
    if (m.perx) d.x = pmod(d.x + nc.x / 2, nc.x) - nc.x / 2;
    d.z = pmod(d.z + nc.z / 2, nc.z) - nc.z / 2;
    if (d.x == off.x && d.y == off.y && d.z == off.z) {
      const int E0 = F.box[0], E1 = F.box[1], E2 = F.box[2];
      const float* boxes = reinterpret_cast<const float*>(sbox);
      // A node three or more nodes from the x and y edges and two from the
      // periodic z seam (no periodic x), on a z axis longer than twice the
      // box: every range below is one box position, or none when that
      // position lies outside the box (the z wrap of a negative position
      // lands past the box); the same terms in the same order.
      if (!m.perx && qc.x >= 3 && qc.x <= nc.x - 4 && qc.y >= 3 && qc.z >= 2 && qc.z <= nc.z - 3 && nc.z >= 2 * E2 + 8) {
        for (int c = 0; c < total; c++) {
          const float w = wsup[c]; if (w == 0) continue;
          const int bx = J[0] - asup[c][0], by = J[1] - asup[c][1], bz = J[2] - asup[c][2];
          if ((unsigned)bx < (unsigned)E0 && (unsigned)by < (unsigned)E1 && (unsigned)bz < (unsigned)E2)
            acc += w * boxes[c * Ep + bx + E0 * (by + E1 * bz)];
        }
        row[(int64_t)k * C.count] = acc;
        return;
      }
      for (int c = 0; c < total; c++) {
        const float w = wsup[c]; if (w == 0) continue;
        const int ax = asup[c][0], ay = asup[c][1], az = asup[c][2];
        const float* T = boxes + c * Ep;
        // box ranges whose positions map to J along each axis
        int ylo = J[1] == 0 ? 0 : J[1] - ay, yhi = J[1] - ay;
        if (ylo < 0) ylo = 0; if (yhi > E1 - 1) yhi = E1 - 1;
        int xlo, xhi, xstep;
        if (m.perx) { xlo = pmod(J[0] - ax, nc.x); xhi = E0 - 1; xstep = nc.x; }
        else { xlo = J[0] == 0 ? 0 : J[0] - ax; xhi = J[0] == nc.x - 1 ? E0 - 1 : J[0] - ax; xstep = 1;
               if (xlo < 0) xlo = 0; if (xhi > E0 - 1) xhi = E0 - 1; }
        for (int bz = pmod(J[2] - az, nc.z); bz < E2; bz += nc.z)
          for (int by = ylo; by <= yhi; by++)
            for (int bx = xlo; bx <= xhi; bx += xstep) acc += w * T[bx + E0 * (by + E1 * bz)];
      }
    }
  }
  row[(int64_t)k * C.count] = acc;
}
// Keep the canonical (upper) entry of every coupling and mirror it, so the
// coarse operator is exactly symmetric; non-canonical duplicate offsets stay 0.
// `lanes` lanes per row, each taking every lanes-th entry: the entries are
// independent (a mirrored entry reads a canonical one, never written here).
// 8 lanes on a level of more than 2^17 rows, 16 below (measured: 16 lanes
// take 19 against 21 us on the 37,440-row level, 40 against 33 us on the
// 297,600-row level)
inline int sym_lanes(int64_t rows) { return rows > (1 << 17) ? 8 : 16; }
// Consecutive threads take consecutive rows of the same entry group, so the
// k-major rows are read and written coalesced; the slot table is in LDS.
__global__ void __launch_bounds__(256) k_mg_symmetrize(Ctx s, MGDev m, int l, int lanes) {
  if (!build_on(s, m, false)) return;
  __shared__ signed char slots[128];
  for (int k = threadIdx.x; k < 125; k += blockDim.x) slots[k] = m.slots[125 * l + k];
  __syncthreads();
  const MGLevel& C = m.lv[l];
  const uint32_t gid = (uint32_t)GID(), n = (uint32_t)C.count;
  const int lane = (int)(gid / n); const int64_t I = gid - (uint32_t)lane * n; if (lane >= lanes) return;
  const I3 nd = mdims(m, l), q = coord(I, nd);
  float* base = m.rows + C.row_off;
  for (int k = lane; k < C.width; k += lanes) {
    I3 off = moffset(m, l, k); int J = mwrap(m, l, mk(q.x + off.x, q.y + off.y, q.z + off.z));
    if (J < 0) { base[(int64_t)k * C.count + I] = 0; continue; }
    // a neighbour reached without a wrap lies at coordinate q + off, so
    // mdelta(I, J) is the minimal image of off and mdelta(J, I) that of -off
    const bool direct = (unsigned)(q.x + off.x) < (unsigned)nd.x && (unsigned)(q.z + off.z) < (unsigned)nd.z;
    I3 d, e;
    if (direct) {
      d = off; e = mk(-off.x, -off.y, -off.z);
      if (m.perx) { d.x = pmod(d.x + nd.x / 2, nd.x) - nd.x / 2; e.x = pmod(e.x + nd.x / 2, nd.x) - nd.x / 2; }
      d.z = pmod(d.z + nd.z / 2, nd.z) - nd.z / 2; e.z = pmod(e.z + nd.z / 2, nd.z) - nd.z / 2;
    } else d = mdelta(m, l, (int)I, J);
    if (d.x != off.x || d.y != off.y || d.z != off.z) { base[(int64_t)k * C.count + I] = 0; continue; }
    if (J < I) {
      if (!direct) e = mdelta(m, l, J, (int)I);
      const int sl = slots[slot125(e)];
      base[(int64_t)k * C.count + I] = base[(int64_t)sl * C.count + J];
    }
  }
}
// pressure_mg_coarse.glslinc:35-49 (op 11): dense Cholesky in LDS, store L^-1
// Dense Cholesky factor of the coarsest level and the inverse of the factor
// (pressure_multigrid.glslinc, coarsest-level solve).  Written right-looking:
// once column k is final every entry below and to the right subtracts its
// product with column k, so each entry receives exactly the subtractions of
// the left-looking loops (a - L_i0 L_k0 - L_i1 L_k1 - ..., same operands,
// same j order) while the updates of one step run on all threads.  Every
// thread takes the square root of the diagonal and the quotients of column k
// it needs itself (the same operations on the same values, so the same bits
// as the stored factor), which leaves one barrier per step; row k - 1 of the
// inverse of the factor, built the same way, shares step k.
constexpr int CHOL_THREADS = 1024;  // sixteen waves: a step's rows spread four to a wave
__global__ void __launch_bounds__(CHOL_THREADS) k_mg_cholesky(Ctx s, MGDev m, int l) {
  if (!build_on(s, m, false)) return;
  // dense: the working matrix (lower triangle); lf: the factor; acc: rows of
  // the inverse still accumulating
  __shared__ float dense[64 * 64], lf[64 * 64], acc[64 * 64];
  // the level's rows (k-major, width x N), active flags and delta slots staged
  // first with independent loads, so the N^2 entries below read LDS only
  __shared__ float rows[125 * 64];
  __shared__ signed char sl125[128];
  __shared__ unsigned char act[64];
  const int lid = threadIdx.x, N = m.lv[l].count, W = m.lv[l].width;
  if (l > 0) for (int k = lid; k < W * N; k += CHOL_THREADS) rows[k] = m.rows[m.lv[l].row_off + k];
  for (int k = lid; k < 125; k += CHOL_THREADS) sl125[k] = m.slots[125 * l + k];
  for (int k = lid; k < N; k += CHOL_THREADS) act[k] = mactive(m, l, k) ? 1 : 0;
  __syncthreads();
  for (int k = lid; k < N * N; k += CHOL_THREADS) {
    int i = k / N, j = k % N; float v;
    if (!act[i] || !act[j]) v = i == j ? 1 : 0;
    else {
      int a = i < j ? i : j, b = i < j ? j : i; I3 d = mdelta(m, l, a, b);
      int sl = in_box2(d) ? sl125[slot125(d)] : -1;
      v = sl < 0 ? 0 : rows[sl * N + a];
    }
    dense[k] = v;
    acc[k] = i == j ? 1 : 0;
  }
  __syncthreads();
  const int wave = lid >> 6, lane = lid & 63;
  for (int k = 0; k <= N; k++) {
    if (k < N) {
      float d = dense[k * N + k];
      const bool bad = !(d > 0) || isnan(d) || isinf(d);
      if (bad) d = 1;
      const float lkk = sqrtf(d);
      if (lid == 0) { if (bad) s.c[7] = 33; lf[k * N + k] = lkk; }
      if (lid > k && lid < N) lf[lid * N + k] = dense[lid * N + k] / lkk;
      const int c = k + 1 + lane;
      if (c < N) {
        const float lc = dense[c * N + k] / lkk;
        for (int r = c + ((wave - (c - k - 1)) & (CHOL_THREADS / 64 - 1)); r < N; r += CHOL_THREADS / 64)
          dense[r * N + c] -= dense[r * N + k] / lkk * lc;
      }
    }
    if (k >= 1 && lane < N) {
      const int row = k - 1;
      const float x = acc[row * N + lane] / lf[row * N + row];
      if (wave == 0) m.linv[row * N + lane] = x;
      for (int r = row + 1 + wave; r < N; r += CHOL_THREADS / 64)
        acc[r * N + lane] -= lf[r * N + row] * x;
    }
    __syncthreads();
  }
}

static inline unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// water/pressure_multigrid.gd:121-128 + water/pressure_hierarchy_state.gd:36-41
void wv_mg_build(const Ctx& s, const MGDev& m, hipStream_t st) {
  hipLaunchKernelGGL(k_audit0, dim3(1), dim3(256), 0, st, s, m);
  hipLaunchKernelGGL(k_audit1, dim3(nblocks(s.N, 256)), dim3(256), 0, st, s, m);
  hipLaunchKernelGGL(k_audit2, dim3(1), dim3(64), 0, st, s, m);
  hipLaunchKernelGGL(k_audit3, dim3(nblocks(s.N, 256)), dim3(256), 0, st, s, m);
  for (int l = 0; l < m.L; l++) hipLaunchKernelGGL(k_mg_classify, dim3(nblocks(m.lv[l].count, 256)), dim3(256), 0, st, s, m, l);
  hipLaunchKernelGGL(k_mg_l1_fine, dim3(nblocks(m.lv[0].count, 256)), dim3(256), 0, st, s, m);
  for (int l = 0; l + 1 < m.L; l++) {
    size_t lds = sizeof(float) * m.lv[l + 1].width * GAL_THREADS;
    int G = m.lv[l + 1].rlanes > 8 ? 8 : m.lv[l + 1].rlanes;
    const size_t ap_lds = sizeof(float) * 128 * m.lv[l].box[0] * m.lv[l].box[1] * m.lv[l].box[2];
    const int ga = m.lv[l].ga;  // lanes per fine row (planned with the sweeps' lane target)
    hipLaunchKernelGGL(k_mg_galerkin_ap, dim3(nblocks((int64_t)m.lv[l].count * ga, 128)), dim3(128), ap_lds, st, s, m, l, ga);
    // wide coarse stencils (125 entries: a block's 128 lanes) gather per entry;
    // narrow ones split each node's fine support over lanes with LDS slices
    const size_t gather_lds = sizeof(float4) * GATHER_SUPPORT * ((m.lv[l].box[0] * m.lv[l].box[1] * m.lv[l].box[2] + 3) / 4);
    if (m.lv[l + 1].width > 64) hipLaunchKernelGGL(k_mg_galerkin_gather, dim3(m.lv[l + 1].count), dim3(128), gather_lds, st, s, m, l);
    else hipLaunchKernelGGL(k_mg_galerkin_ptap, dim3(nblocks((int64_t)m.lv[l + 1].count * G, GAL_THREADS)), dim3(GAL_THREADS), lds, st, s, m, l);
    const int sl = sym_lanes(m.lv[l + 1].count);
    hipLaunchKernelGGL(k_mg_symmetrize, dim3(nblocks((int64_t)m.lv[l + 1].count * sl, 256)), dim3(256), 0, st, s, m, l + 1, sl);
    hipLaunchKernelGGL(k_mg_l1_coarse, dim3(nblocks(m.lv[l + 1].count, 256)), dim3(256), 0, st, s, m, l + 1);
  }
  if (m.lv[m.L - 1].count <= 64) hipLaunchKernelGGL(k_mg_cholesky, dim3(1), dim3(CHOL_THREADS), 0, st, s, m, m.L - 1);
  hipLaunchKernelGGL(k_audit4, dim3(1), dim3(1), 0, st, s, m);
}

// ---- the smallest levels in one workgroup -----------------------------------
// Below some level T the levels are a few hundred nodes: one launch per sweep
// costs more than the sweep.  k_mg_tail runs cycle(T) -- the sweeps, residual
// and restriction of each level on the way down, the dense coarse solve, the
// prolongation and post-sweeps on the way up -- in one 1024-thread workgroup
// with the levels' rows, vectors, stencil offsets and L^-1 in LDS, a barrier
// between phases.  It mirrors the launch sequence phase for phase (one thread
// per row, each row's sum in stencil order) and writes every vector field of
// the tail levels back, so the level vectors end as the launches leave them
// up to the rounding of the row sums' order.
#define TAIL_THREADS 1024
struct TailLayout { int T, L, coef[WV_MAXL], vec[WV_MAXL], off[WV_MAXL], linv, rhs, floats; };
static TailLayout tail_layout(const MGDev& m, int T) {
  TailLayout t; t.T = T; t.L = m.L; int at = 0;
  for (int l = T; l < m.L; l++) {
    const MGLevel& L = m.lv[l];
    t.coef[l] = at; at += L.width * L.count;
    t.vec[l] = at; at += 8 * L.count;
    t.off[l] = at; at += 4 * L.width;  // dx, dy, dz, linear delta per entry
  }
  int nl = m.lv[m.L - 1].count;
  t.linv = at; at += nl * nl; t.rhs = at; at += 64;
  t.floats = at;
  return t;
}
static int tail_lds_limit() {
  static int limit = -1;
  if (limit < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
    limit = v;
  }
  return limit;
}
// first tail level, or -1: the tail holds at least two levels, ends in the
// dense coarse solve, and fits the workgroup's LDS
static int tail_start(const MGDev& m) {
  if (m.L < 3 || m.lv[m.L - 1].count > 64) return -1;
  int best = -1;
  for (int T = m.L - 2; T >= 1; T--) {
    if ((size_t)tail_layout(m, T).floats * sizeof(float) > (size_t)tail_lds_limit()) break;
    best = T;
  }
  return best;
}
int wv_mg_tail_level(const MGDev& m) { return tail_start(m); }

// dst[e] = src(e) for e < n over the workgroup, eight elements per thread in
// flight (all loads of a round before its stores)
template <class Src>
__device__ __forceinline__ void tail_copy(float* dst, int n, Src src) {
  constexpr int U = 8;
  for (int base = threadIdx.x; base < n; base += U * TAIL_THREADS) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; u++) { const int e = base + u * TAIL_THREADS; v[u] = src(e < n ? e : n - 1); }  // in-range loads, unconditional
#pragma unroll
    for (int u = 0; u < U; u++) { const int e = base + u * TAIL_THREADS; if (e < n) dst[e] = v[u]; }
  }
}
template <bool H>
__global__ void __launch_bounds__(TAIL_THREADS) k_mg_tail(Ctx s, MGDev m, TailLayout tl, int phase, int smoothing) {
  if (!mg_enabled(s, phase)) return;
  extern __shared__ float sm[];
  const int tid = threadIdx.x;
  auto vec = [&](int l, int field) { return sm + tl.vec[l] + field * m.lv[l].count; };
  auto active = [&](int l, int i) { return i >= 0 && vec(l, 3)[i] > .5f; };
  for (int l = tl.T; l < tl.L; l++) {
    const MGLevel& L = m.lv[l]; const int n = L.count, W = L.width;
    const int64_t ro = L.row_off, vo = L.vec_off;
    if (H) tail_copy(sm + tl.coef[l], W * n, [&](int e) { return bf16_to_f32(m.rows16[ro + e]); });
    else tail_copy(sm + tl.coef[l], W * n, [&](int e) { return m.rows[ro + e]; });
    tail_copy(sm + tl.vec[l], 8 * n, [&](int e) { return m.vec[vo + e]; });
    for (int k = tid; k < W; k += TAIL_THREADS) {
      int* o = reinterpret_cast<int*>(sm) + tl.off[l] + 4 * k;
      o[0] = m.offsets[3 * (L.off_start + k)]; o[1] = m.offsets[3 * (L.off_start + k) + 1];
      o[2] = m.offsets[3 * (L.off_start + k) + 2]; o[3] = m.deltas[L.off_start + k];
    }
  }
  const int nl = m.lv[tl.L - 1].count;
  tail_copy(sm + tl.linv, nl * nl, [&](int e) { return m.linv[e]; });
  __syncthreads();
  // A x_source at row i of level l, entries in stencil order (coarse_apply;
  // on wrap1 levels the neighbour is row + delta with one periodic correction)
  auto apply = [&](int l, int source, int i) {
    const MGLevel& L = m.lv[l]; const int n = L.count, W = L.width;
    const int* off = reinterpret_cast<const int*>(sm) + tl.off[l];
    const float* A = sm + tl.coef[l]; const float* x = vec(l, source);
    I3 q = coord(i, mdims(m, l)); float sum = 0;
    if (L.wrap1) {
      const int nz = L.n[2], nx = L.n[0];
#pragma unroll 9
      for (int k = 0; k < W; k++) {
        int j = i + off[4 * k + 3], zz = q.z + off[4 * k + 2];
        j += zz < 0 ? n : (zz >= nz ? -n : 0);
        if (m.perx) { int xx = q.x + off[4 * k]; j += xx < 0 ? nx : (xx >= nx ? -nx : 0); }
        j = min(max(j, 0), n - 1);
        sum += A[k * n + i] * x[j];
      }
    } else {
      for (int k = 0; k < W; k++) {
        int j = mwrap(m, l, mk(q.x + off[4 * k], q.y + off[4 * k + 1], q.z + off[4 * k + 2]));
        sum += A[k * n + i] * x[j >= 0 ? j : 0];
      }
    }
    return sum;
  };
  // coarse_row OP 0 (Chebyshev sweep k) and OP 1 (residual)
  auto sweep = [&](int l, int source, int k) {
    const int n = m.lv[l].count;
    for (int i = tid; i < n; i += TAIL_THREADS) {
      float ax = apply(l, source, i), xs = vec(l, source)[i];
      if (!active(l, i)) ax = xs;
      float r = vec(l, 0)[i] - ax;
      float d = cheb_dir(k, vec(l, 7)[i], vec(l, 4)[i], r);
      vec(l, 7)[i] = d; vec(l, 3 - source)[i] = xs + d;
    }
    __syncthreads();
  };
  auto residual = [&](int l, int field) {
    const int n = m.lv[l].count;
    for (int i = tid; i < n; i += TAIL_THREADS) {
      float ax = apply(l, field, i), xs = vec(l, field)[i];
      if (!active(l, i)) ax = xs;
      vec(l, 6)[i] = vec(l, 0)[i] - ax;
    }
    __syncthreads();
  };
  // k_mg_restrict_coop into level lc, children in order on one thread
  auto restrict_to = [&](int lc) {
    const MGLevel& C = m.lv[lc]; const int parent = lc - 1; const MGLevel& P = m.lv[parent];
    I3 nf = mdims(m, parent), nlc = mdims(m, lc);
    for (int I = tid; I < C.count; I += TAIL_THREADS) {
      float sum = 0;
      if (active(lc, I)) {
        I3 qc = coord(I, nlc);
        int lo[3], ext[3] = {1, 1, 1}, nfa[3] = {nf.x, nf.y, nf.z}, nla[3] = {nlc.x, nlc.y, nlc.z}, qca[3] = {qc.x, qc.y, qc.z};
        for (int a = 0; a < 3; a++) lo[a] = qca[a] * P.f[a];
        if (P.aggregate) { ext[0] = P.f[0]; ext[1] = P.f[1]; ext[2] = P.f[2]; }
        else for (int a = 0; a < 3; a++) if (P.f[a] > 1) {
          lo[a]--; ext[a] = 4;
          if ((a == 2 || (a == 0 && m.perx)) && nfa[a] <= 4) { lo[a] = 0; ext[a] = nfa[a]; }
        }
        int total = ext[0] * ext[1] * ext[2];
        for (int t = 0; t < total; t++) {
          int x = t % ext[0], y = (t / ext[0]) % ext[1], z = t / (ext[0] * ext[1]);
          int j = mwrap(m, parent, mk(lo[0] + x, lo[1] + y, lo[2] + z)); if (!active(parent, j)) continue;
          float weight = 1;
          if (!P.aggregate) {
            I3 q = coord(j, nf); int qa[3] = {q.x, q.y, q.z};
            for (int a = 0; a < 3; a++) {
              float pos = lattice_pos(qa[a], P.f[a]); int base = (int)floorf(pos); float tt = fract(pos);
              float w = 0;
              for (int bit = 0; bit < P.f[a]; bit++) {
                int nb = base + bit;
                if (a == 2 || (a == 0 && m.perx)) nb = pmod(nb, nla[a]);
                else if (a == 0) nb = iclamp(nb, 0, nla[a] - 1); else nb = nb > 0 ? nb : 0;
                if (nb == qca[a]) w += P.f[a] == 1 ? 1 : (bit == 1 ? tt : 1 - tt);
              }
              weight *= w;
            }
          }
          sum += weight * vec(parent, 6)[j];
        }
      }
      vec(lc, 0)[I] = sum; vec(lc, 1)[I] = 0;
      if (lc < tl.L - 1) cheb_first(&vec(lc, 7)[I], &vec(lc, 2)[I], vec(lc, 4)[I], sum);
      else vec(lc, 2)[I] = 0;
    }
    __syncthreads();
  };
  // prolong_row with the transfers of mtransfers, active flags from LDS
  auto prolong = [&](int l, int ff, int cf) {
    const MGLevel& L = m.lv[l]; I3 n1 = mdims(m, l + 1), nlv = mdims(m, l);
    for (int i = tid; i < L.count; i += TAIL_THREADS) {
      float add = 0;
      if (active(l, i)) {
        I3 q = coord(i, nlv);
        if (L.aggregate) {
          int j = (int)index3(mk(fdiv(q.x, L.f[0]), fdiv(q.y, L.f[1]), fdiv(q.z, L.f[2])), n1);
          if (active(l + 1, j)) add += 1 * vec(l + 1, cf)[j];
        } else {
          int base[3]; float t[3]; int qq[3] = {q.x, q.y, q.z};
          for (int a = 0; a < 3; a++) { float pos = lattice_pos(qq[a], L.f[a]); base[a] = (int)floorf(pos); t[a] = fract(pos); }
          for (int z = 0; z < L.f[2]; z++) for (int y = 0; y < L.f[1]; y++) for (int x = 0; x < L.f[0]; x++) {
            int bit[3] = {x, y, z}; I3 nb = mk(base[0] + x, base[1] + y, base[2] + z); float ww = 1;
            for (int a = 0; a < 3; a++) if (L.f[a] > 1) ww *= bit[a] == 1 ? t[a] : 1 - t[a];
            if (nb.y >= n1.y) continue; if (nb.y < 0) nb.y = 0;
            nb.x = m.perx ? pmod(nb.x, n1.x) : iclamp(nb.x, 0, n1.x - 1); nb.z = pmod(nb.z, n1.z);
            int j = (int)index3(nb, n1); if (!active(l + 1, j) || ww == 0) continue;
            add += ww * vec(l + 1, cf)[j];
          }
        }
      }
      vec(l, ff)[i] += add;
    }
    __syncthreads();
  };
  // every level here is smoothed (the last one is solved), so the kernel that
  // set its b -- the restriction launch into level T, restrict_to below it --
  // did its first sweep (cheb_first)
  int fields[WV_MAXL];
  for (int l = tl.T; l < tl.L - 1; l++) {
    int f = 2;
    for (int k = 2; k <= smoothing; k++) { sweep(l, f, k); f = 3 - f; }
    fields[l] = f;
    residual(l, f);
    restrict_to(l + 1);
  }
  {  // k_mg_coarse_solve on the last level: x = L^-T (L^-1 b), field 1
    const float* Li = sm + tl.linv; float* rhs = sm + tl.rhs; const int N = nl, lc = tl.L - 1;
    if (tid < N) { float d = 0; for (int j = 0; j <= tid; j++) d += Li[tid * N + j] * vec(lc, 0)[j]; rhs[tid] = d; }
    __syncthreads();
    if (tid < N) { float d = 0; for (int j = tid; j < N; j++) d += Li[j * N + tid] * rhs[j]; vec(lc, 1)[tid] = d; }
    __syncthreads();
  }
  int cf = 1;
  for (int l = tl.L - 2; l >= tl.T; l--) {
    int f = fields[l];
    prolong(l, f, cf);
    for (int k = 1; k <= smoothing; k++) { sweep(l, f, k); f = 3 - f; }
    cf = f;
  }
  for (int l = tl.T; l < tl.L; l++) {
    const MGLevel& L = m.lv[l];
    tail_copy(m.vec + L.vec_off, 8 * L.count, [&](int e) { return sm[tl.vec[l] + e]; });
  }
}
// the field holding level T's result after the tail (the host's replay of the
// field alternation above)
static int tail_field(int smoothing) {
  int down = smoothing % 2 ? 2 : 1;
  return smoothing % 2 ? 3 - down : down;
}

// ---- the tail with index tables ---------------------------------------------
// When every transfer inside the tail aggregates (piecewise-constant, the
// planner's levels >= 2) and its levels have fewer than 32768 nodes, the
// neighbour index each row entry reads (the wrapped or clamped index of
// k_mg_tail's apply), each coarse node's children and each fine node's parent
// depend on the grid only.  They are tabulated once per solver as 16-bit
// indices (k_mg_tail_tables) and loaded into LDS with the rows, so a phase is
// loads and the row sums in stencil order, with no index arithmetic.  Only
// the smoothed levels' rows are loaded (the last level is solved with L^-1),
// all loads are 16 bytes wide, and the coarse solve's products are taken by
// the whole workgroup before each row's serial sum.  Operations, operands and
// orders are those of k_mg_tail.
struct TailFast {
  int T, L, words, tab, tab_words, linv, rhs, diag;
  int coef[WV_MAXL], vec[WV_MAXL];            // word offsets in LDS
  int J[WV_MAXL], ch[WV_MAXL], par[WV_MAXL];  // 16-bit entry offsets in the table region
  int CH[WV_MAXL];                            // children per node of level l (restriction into l)
};
static bool tail_fast_layout(const MGDev& m, int T, TailFast& t) {
  if (T < 1 || m.L - T < 2 || m.lv[m.L - 1].count > 64) return false;
  for (int l = T; l < m.L; l++) if (m.lv[l].count >= 32768) return false;
  for (int l = T; l < m.L - 1; l++) if (!m.lv[l].aggregate) return false;
  t = TailFast{};
  t.T = T; t.L = m.L;
  auto up4 = [](int x) { return (x + 3) & ~3; };
  int at = 0, sh = 0;
  for (int l = T; l < m.L - 1; l++) { t.coef[l] = at; at = up4(at + m.lv[l].width * m.lv[l].count); }
  for (int l = T; l < m.L; l++) { t.vec[l] = at; at = up4(at + 8 * m.lv[l].count); }
  for (int l = T; l < m.L - 1; l++) { t.J[l] = sh; sh += m.lv[l].width * m.lv[l].count; }
  for (int l = T + 1; l < m.L; l++) {
    const MGLevel& P = m.lv[l - 1];
    t.CH[l] = P.f[0] * P.f[1] * P.f[2]; t.ch[l] = sh; sh += t.CH[l] * m.lv[l].count;
  }
  for (int l = T; l < m.L - 1; l++) { t.par[l] = sh; sh += m.lv[l].count; }
  sh = (sh + 7) & ~7;                  // whole 16-byte quads
  t.tab = at; t.tab_words = sh / 2; at = up4(at + t.tab_words);
  const int N = m.lv[m.L - 1].count;
  t.linv = at; at = up4(at + N * N);
  t.rhs = at; at += 64; t.diag = at; at += 64;
  t.words = at;
  return (size_t)t.words * sizeof(float) <= (size_t)tail_lds_limit();
}
// the table region's 16-bit entries (0 when the tail has no table form)
int64_t wv_mg_tail_table_shorts(const MGDev& m) {
  TailFast t;
  return m.tail >= 1 && tail_fast_layout(m, m.tail, t) ? 2 * (int64_t)t.tab_words : 0;
}
__global__ void __launch_bounds__(256) k_mg_tail_tables(MGDev m, TailFast t, int16_t* tab) {
  const int total = 2 * t.tab_words;
  for (int e = (int)GID(); e < total; e += gridDim.x * blockDim.x) {
    int v = -1;
    for (int l = t.T; l < t.L - 1; l++) {          // row entries: k-major, the index apply reads
      const MGLevel& L = m.lv[l]; const int n = L.count;
      if (e < t.J[l] || e >= t.J[l] + L.width * n) continue;
      const int k = (e - t.J[l]) / n, i = (e - t.J[l]) - k * n;
      const I3 q = coord(i, mdims(m, l));
      const int* o = m.offsets + 3 * (L.off_start + k);
      if (L.wrap1) {
        int j = i + m.deltas[L.off_start + k], zz = q.z + o[2];
        j += zz < 0 ? n : (zz >= L.n[2] ? -n : 0);
        if (m.perx) { int xx = q.x + o[0]; j += xx < 0 ? L.n[0] : (xx >= L.n[0] ? -L.n[0] : 0); }
        v = min(max(j, 0), n - 1);
      } else {
        const int j = mwrap(m, l, mk(q.x + o[0], q.y + o[1], q.z + o[2]));
        v = j >= 0 ? j : 0;
      }
    }
    for (int l = t.T + 1; l < t.L; l++) {          // children of a coarse node: child-major, in child order
      const int n = m.lv[l].count;
      if (e < t.ch[l] || e >= t.ch[l] + t.CH[l] * n) continue;
      const MGLevel& P = m.lv[l - 1];
      const int c = (e - t.ch[l]) / n, I = (e - t.ch[l]) - c * n;
      const I3 qc = coord(I, mdims(m, l));
      const int x = c % P.f[0], y = (c / P.f[0]) % P.f[1], z = c / (P.f[0] * P.f[1]);
      v = mwrap(m, l - 1, mk(qc.x * P.f[0] + x, qc.y * P.f[1] + y, qc.z * P.f[2] + z));
    }
    for (int l = t.T; l < t.L - 1; l++) {          // parent of a fine node
      const MGLevel& L = m.lv[l];
      if (e < t.par[l] || e >= t.par[l] + L.count) continue;
      const I3 q = coord(e - t.par[l], mdims(m, l));
      v = (int)index3(mk(fdiv(q.x, L.f[0]), fdiv(q.y, L.f[1]), fdiv(q.z, L.f[2])), mdims(m, l + 1));
    }
    tab[e] = (int16_t)v;
  }
}
void wv_mg_tail_tables(const MGDev& m, int16_t* tab, hipStream_t st) {
  TailFast t;
  if (!tab || m.tail < 1 || !tail_fast_layout(m, m.tail, t)) return;
  hipLaunchKernelGGL(k_mg_tail_tables, dim3(64), dim3(256), 0, st, m, t, tab);
}
// n words from global to LDS, 16 bytes a load where the source allows, every
// load of a thread's round issued before its stores
__device__ __forceinline__ void tail_load_words(float* dst, const float* src, int n) {
  if (((uintptr_t)src & 15) == 0) {
    const int n4 = n >> 2; const float4* s4 = reinterpret_cast<const float4*>(src); float4* d4 = reinterpret_cast<float4*>(dst);
    constexpr int U = 4;
    for (int base = threadIdx.x; base < n4; base += U * TAIL_THREADS) {
      float4 v[U];
#pragma unroll
      for (int u = 0; u < U; u++) { const int e = base + u * TAIL_THREADS; v[u] = s4[e < n4 ? e : n4 - 1]; }
#pragma unroll
      for (int u = 0; u < U; u++) { const int e = base + u * TAIL_THREADS; if (e < n4) d4[e] = v[u]; }
    }
    for (int e = 4 * n4 + (int)threadIdx.x; e < n; e += TAIL_THREADS) dst[e] = src[e];
  } else {
    tail_copy(dst, n, [&](int e) { return src[e]; });
  }
}
__device__ __forceinline__ void tail_load_bf16(float* dst, const uint16_t* src, int n) {
  if (((uintptr_t)src & 7) == 0) {
    const int n4 = n >> 2; const uint2* s4 = reinterpret_cast<const uint2*>(src); float4* d4 = reinterpret_cast<float4*>(dst);
    constexpr int U = 4;
    for (int base = threadIdx.x; base < n4; base += U * TAIL_THREADS) {
      uint2 v[U];
#pragma unroll
      for (int u = 0; u < U; u++) { const int e = base + u * TAIL_THREADS; v[u] = s4[e < n4 ? e : n4 - 1]; }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int e = base + u * TAIL_THREADS;
        if (e < n4) d4[e] = make_float4(__uint_as_float(v[u].x << 16), __uint_as_float(v[u].x & 0xFFFF0000u),
                                        __uint_as_float(v[u].y << 16), __uint_as_float(v[u].y & 0xFFFF0000u));
      }
    }
    for (int e = 4 * n4 + (int)threadIdx.x; e < n; e += TAIL_THREADS) dst[e] = bf16_to_f32(src[e]);
  } else {
    tail_copy(dst, n, [&](int e) { return bf16_to_f32(src[e]); });
  }
}
__device__ __forceinline__ void tail_store_words(float* dst, const float* src, int n) {
  if (((uintptr_t)dst & 15) == 0) {
    const int n4 = n >> 2; const float4* s4 = reinterpret_cast<const float4*>(src); float4* d4 = reinterpret_cast<float4*>(dst);
    for (int e = threadIdx.x; e < n4; e += TAIL_THREADS) d4[e] = s4[e];
    for (int e = 4 * n4 + (int)threadIdx.x; e < n; e += TAIL_THREADS) dst[e] = src[e];
  } else {
    for (int e = threadIdx.x; e < n; e += TAIL_THREADS) dst[e] = src[e];
  }
}
template <bool H, int WC>
__global__ void __launch_bounds__(TAIL_THREADS) k_mg_tail_fast(Ctx s, MGDev m, TailFast tl, int phase, int smoothing) {
  if (!mg_enabled(s, phase)) return;
  extern __shared__ float sm[];
  const int tid = threadIdx.x;
  const int lc = tl.L - 1, N = m.lv[lc].count;
  for (int l = tl.T; l < tl.L - 1; l++) {
    const MGLevel& L = m.lv[l];
    if (H) tail_load_bf16(sm + tl.coef[l], m.rows16 + L.row_off, L.width * L.count);
    else tail_load_words(sm + tl.coef[l], m.rows + L.row_off, L.width * L.count);
  }
  for (int l = tl.T; l < tl.L; l++) tail_load_words(sm + tl.vec[l], m.vec + m.lv[l].vec_off, 8 * m.lv[l].count);
  tail_load_words(sm + tl.tab, reinterpret_cast<const float*>(m.ttab), tl.tab_words);
  tail_load_words(sm + tl.linv, m.linv, N * N);
  __syncthreads();
  const int16_t* tab = reinterpret_cast<const int16_t*>(sm + tl.tab);
  auto vec = [&](int l, int field) { return sm + tl.vec[l] + field * m.lv[l].count; };
  auto active = [&](int l, int i) { return vec(l, 3)[i] > .5f; };
  // sum_k A[i, k] x[J[i, k]] in stencil order
  auto apply = [&](int l, int source, int i) {
    const int n = m.lv[l].count;
    const float* A = sm + tl.coef[l]; const int16_t* J = tab + tl.J[l]; const float* x = vec(l, source);
    float sum = 0;
    if (WC > 0) {
      float a[WC > 0 ? WC : 1]; int j[WC > 0 ? WC : 1];
#pragma unroll
      for (int k = 0; k < WC; k++) { a[k] = A[k * n + i]; j[k] = J[k * n + i]; }
#pragma unroll
      for (int k = 0; k < WC; k++) a[k] = a[k] * x[j[k]];
#pragma unroll
      for (int k = 0; k < WC; k++) sum += a[k];
    } else {
      const int W = m.lv[l].width;
#pragma unroll 9
      for (int k = 0; k < W; k++) sum += A[k * n + i] * x[J[k * n + i]];
    }
    return sum;
  };
  auto sweep = [&](int l, int source, int k) {
    const int n = m.lv[l].count;
    for (int i = tid; i < n; i += TAIL_THREADS) {
      const float xs = vec(l, source)[i], b = vec(l, 0)[i], dprev = vec(l, 7)[i], dinv = vec(l, 4)[i];
      const bool on = active(l, i);
      float ax = apply(l, source, i);
      if (!on) ax = xs;
      const float d = cheb_dir(k, dprev, dinv, b - ax);
      vec(l, 7)[i] = d; vec(l, 3 - source)[i] = xs + d;
    }
    __syncthreads();
  };
  auto residual = [&](int l, int field) {
    const int n = m.lv[l].count;
    for (int i = tid; i < n; i += TAIL_THREADS) {
      const float xs = vec(l, field)[i], b = vec(l, 0)[i];
      const bool on = active(l, i);
      float ax = apply(l, field, i);
      if (!on) ax = xs;
      vec(l, 6)[i] = b - ax;
    }
    __syncthreads();
  };
  // aggregation into level c: the active children's residuals in child order
  auto restrict_to = [&](int c) {
    const int n = m.lv[c].count, CH = tl.CH[c]; const int16_t* ch = tab + tl.ch[c];
    for (int I = tid; I < n; I += TAIL_THREADS) {
      float sum = 0;
      if (active(c, I)) {
        for (int t = 0; t < CH; t++) {
          const int j = ch[t * n + I];
          if (j < 0 || !active(c - 1, j)) continue;
          sum += 1.0f * vec(c - 1, 6)[j];
        }
      }
      vec(c, 0)[I] = sum; vec(c, 1)[I] = 0;
      if (c < tl.L - 1) cheb_first(&vec(c, 7)[I], &vec(c, 2)[I], vec(c, 4)[I], sum);
      else vec(c, 2)[I] = 0;
    }
    __syncthreads();
  };
  auto prolong = [&](int l, int ff, int cf) {
    const int n = m.lv[l].count; const int16_t* par = tab + tl.par[l];
    for (int i = tid; i < n; i += TAIL_THREADS) {
      float add = 0;
      if (active(l, i)) { const int j = par[i]; if (active(l + 1, j)) add += 1 * vec(l + 1, cf)[j]; }
      vec(l, ff)[i] += add;
    }
    __syncthreads();
  };
  int fields[WV_MAXL];
  for (int l = tl.T; l < tl.L - 1; l++) {
    int f = 2;   // sweep 1 done where b was set (cheb_first)
    for (int k = 2; k <= smoothing; k++) { sweep(l, f, k); f = 3 - f; }
    fields[l] = f;
    residual(l, f);
    restrict_to(l + 1);
  }
  {  // x = L^-T (L^-1 b) on the last level: the products first, then each row's sum in order
    float* Li = sm + tl.linv; float* rhs = sm + tl.rhs; float* dg = sm + tl.diag; const float* b = vec(lc, 0);
    for (int e = tid; e < N * N; e += TAIL_THREADS) {
      const int t = e / N, j = e - t * N;   // L^-1 row t, column j: the upper triangle takes the products
      if (j < t) Li[j * N + t] = Li[t * N + j] * b[j];
      else if (j == t) dg[t] = Li[t * N + t] * b[t];
    }
    __syncthreads();
    // each row's products added in order, eight loads ahead of their additions
    if (tid < N) {
      float d = 0;
      for (int j0 = 0; j0 < tid; j0 += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = Li[(j0 + u < tid ? j0 + u : 0) * N + tid];
#pragma unroll
        for (int u = 0; u < 8; u++) if (j0 + u < tid) d += v[u];
      }
      d += dg[tid]; rhs[tid] = d;
    }
    __syncthreads();
    for (int e = tid; e < N * N; e += TAIL_THREADS) {
      const int r = e / N, t = e - r * N;   // entry (r, t) with r >= t: L^-1[r, t] rhs[r], in place
      if (r >= t) Li[r * N + t] = Li[r * N + t] * rhs[r];
    }
    __syncthreads();
    if (tid < N) {
      float d = 0;
      for (int j0 = tid; j0 < N; j0 += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = Li[(j0 + u < N ? j0 + u : tid) * N + tid];
#pragma unroll
        for (int u = 0; u < 8; u++) if (j0 + u < N) d += v[u];
      }
      vec(lc, 1)[tid] = d;
    }
    __syncthreads();
  }
  int cf = 1;
  for (int l = tl.L - 2; l >= tl.T; l--) {
    int f = fields[l];
    prolong(l, f, cf);
    for (int k = 1; k <= smoothing; k++) { sweep(l, f, k); f = 3 - f; }
    cf = f;
  }
  for (int l = tl.T; l < tl.L; l++) tail_store_words(m.vec + m.lv[l].vec_off, sm + tl.vec[l], 8 * m.lv[l].count);
}

// dynamic LDS of k_mg_coarse: the rows' product slices when G > 1 lanes share a row
static size_t coarse_lds(const MGDev& m, int l) {
  const MGLevel& L = m.lv[l];
  return L.lanes > 1 ? sizeof(float) * (256 / L.lanes) * L.width : 0;
}
// a k_mg_coarse sweep, with the compile-time width where the level allows it
template <int OP, bool H>
static void launch_coarse(const Ctx& s, const MGDev& m, int l, int source, int k, int phase, hipStream_t st) {
  const MGLevel& L = m.lv[l];
  const dim3 grid(nblocks((int64_t)L.count * L.lanes, 256));
  // (a compile-time width of 39 with every load of a row issued up front was
  // measured slower on level 1: 13.2 -> 15.0 us, occupancy 8 -> 4 waves)
  hipLaunchKernelGGL((k_mg_coarse<OP, H, 0>), grid, dim3(256), coarse_lds(m, l), st, s, m, l, source, k, phase);
}

template <bool H>
struct CycleBuilder {
  const Ctx& s; const MGDev& m; int phase, smoothing; hipStream_t st;
  unsigned coop_blocks(int l) { return nblocks((int64_t)m.lv[l].count * m.lv[l].lanes, 256); }
  // sweeps k0..smoothing of level l from `field`; returns the field holding x
  int smooth(int l, int field, int k0) {
    for (int k = k0; k <= smoothing; k++) {
      if (l == 0) hipLaunchKernelGGL(k_mg_cheb<H>, dim3(nblocks(m.lv[l].count, 256)), dim3(256), 0, st, s, m, l, field, k, phase);
      else if (m.lv[l].width > 64) hipLaunchKernelGGL((k_mg_coarse_wide<0, H, 8>), dim3(nblocks(m.lv[l].count, 64)), dim3(512), 0, st, s, m, l, field, k, phase);
      else launch_coarse<0, H>(s, m, l, field, k, phase, st);
      field = 3 - field;
    }
    return field;
  }
  // water/pressure_multigrid.gd:141-154
  int cycle(int l) {
    if (l > 0 && l == m.tail) {
      TailFast tf;
      if (m.ttab && tail_fast_layout(m, l, tf)) {
        bool w27 = true;
        for (int k = l; k < m.L - 1; k++) w27 = w27 && m.lv[k].width == 27;
        if (w27) hipLaunchKernelGGL((k_mg_tail_fast<H, 27>), dim3(1), dim3(TAIL_THREADS), tf.words * sizeof(float), st, s, m, tf, phase, smoothing);
        else hipLaunchKernelGGL((k_mg_tail_fast<H, 0>), dim3(1), dim3(TAIL_THREADS), tf.words * sizeof(float), st, s, m, tf, phase, smoothing);
        return tail_field(smoothing);
      }
      TailLayout tl = tail_layout(m, l);
      hipLaunchKernelGGL(k_mg_tail<H>, dim3(1), dim3(TAIL_THREADS), tl.floats * sizeof(float), st, s, m, tl, phase, smoothing);
      return tail_field(smoothing);
    }
    if (l == m.L - 1) {
      if (m.lv[l].count <= 64) { hipLaunchKernelGGL(k_mg_coarse_solve, dim3(1), dim3(64), 0, st, s, m, l, phase); return 1; }
      int tf = 1;
      for (int sweep = 0; sweep < 32; sweep++) {
        if (sweep == 0) hipLaunchKernelGGL(k_mg_jacobi, dim3(nblocks(m.lv[l].count, 256)), dim3(256), 0, st, s, m, l, tf, 1, phase);
        else launch_coarse<2, H>(s, m, l, tf, 0, phase, st);
        tf = 3 - tf;
      }
      return tf;
    }
    // sweep 1 was done by the kernel that set this level's b: the V-cycle's
    // load or the PCG update (level 0), the restriction (below)
    int field = smooth(l, 2, 2);
    if (l == 0) hipLaunchKernelGGL(k_mg_residual<H>, dim3(nblocks(m.lv[l].count, 256)), dim3(256), 0, st, s, m, l, field, phase);
    else if (m.lv[l].width > 64) hipLaunchKernelGGL((k_mg_coarse_wide<1, H, 8>), dim3(nblocks(m.lv[l].count, 64)), dim3(512), 0, st, s, m, l, field, 0, phase);
    else launch_coarse<1, H>(s, m, l, field, 0, phase, st);
    hipLaunchKernelGGL(k_mg_restrict_coop, dim3(nblocks((int64_t)m.lv[l + 1].count * m.lv[l + 1].rlanes, 256)), dim3(256), restrict_lds(m, l + 1), st, s, m, l + 1, phase,
                       (int)first_sweep_fused(m, l + 1));
    int cf = cycle(l + 1);
    hipLaunchKernelGGL(k_mg_prolong, dim3(nblocks(m.lv[l].count, 256)), dim3(256), 0, st, s, m, l, field, cf, phase);
    return smooth(l, field, 1);
  }
};

void wv_mg_cycle(const Ctx& s, const MGDev& m, int phase, int smoothing, bool load, hipStream_t st) {
  if (load) hipLaunchKernelGGL(k_mg_load, dim3(nblocks(m.lv[0].count, 256)), dim3(256), 0, st, s, m, phase, (int)first_sweep_fused(m, 0));
  CycleBuilder<false> b{s, m, phase, smoothing, st};
  int final_field = b.cycle(0);
  hipLaunchKernelGGL(k_mg_out<false>, dim3(nblocks(s.N, WV_RED)), dim3(WV_RED), 0, st, s, m, final_field, phase, 0);
  hipLaunchKernelGGL(k_mg_out<true>, dim3(1), dim3(WV_RED), 0, st, s, m, final_field, phase, 1);
}

// op 3's partials + op 4 when the V-cycle is predicated off
void wv_launch_pcg_init_partials(const Ctx& s, hipStream_t st) {
  hipLaunchKernelGGL(k_pcg_init_partials<false>, dim3(nblocks(s.N, WV_RED)), dim3(WV_RED), 0, st, s, 0);
  hipLaunchKernelGGL(k_pcg_init_partials<true>, dim3(1), dim3(WV_RED), 0, st, s, 1);
}

// Hot kernels launched on their own for live timing (the kernels' outputs are
// V-cycle / PCG scratch that every use recomputes first); phase 3 = ungated.
__global__ void __launch_bounds__(256) k_mg_cheb_timed(Ctx s, MGDev m) {
  int64_t i = GID(); if (i >= m.lv[0].count) return;
  int l = 0, source = 1, k = 2;
  float r = m.vec[at(m, l, 0, i)] - mapply0<false>(m, source, (int)i);
  float d = (float)(2 * k - 3) / (float)(2 * k + 1) * m.vec[at(m, l, 7, i)] + (float)(8 * k - 4) / (float)(2 * k + 1) * m.vec[at(m, l, 4, i)] * r;
  m.vec[at(m, l, 7, i)] = d; m.vec[at(m, l, 2, i)] = m.vec[at(m, l, 1, i)] + d;
}
__global__ void __launch_bounds__(256) k_mg_residual_timed(Ctx s, MGDev m) {
  int64_t i = GID(); if (i >= m.lv[0].count) return;
  m.vec[at(m, 0, 6, i)] = m.vec[at(m, 0, 0, i)] - mapply0<false>(m, 1, (int)i);
}
// plain copy of N floats (the PCG scratch A p into the Galerkin scratch): the
// bandwidth reference for counters and the roofline
__global__ void __launch_bounds__(256) k_copy_timed(const float* __restrict__ src, float* __restrict__ dst, int64_t n) {
  int64_t i = GID(); if (i < n) dst[i] = src[i];
}
// streaming copy of n floats, 16 B per lane, grid-stride: the large-working-set
// bandwidth reference (one warm-up launch, then `reps` timed with events)
__global__ void __launch_bounds__(256) k_copy_stream(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n4) {
  for (int64_t i = GID(); i < n4; i += (int64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}
int wv_time_copy(float* a, float* b, int64_t n, int reps, hipStream_t st, float* ms) {
  const int64_t n4 = n / 4;
  const unsigned grid = 256 * 64;   // 64 blocks per CU
  hipLaunchKernelGGL(k_copy_stream, dim3(grid), dim3(256), 0, st, (const float4*)a, (float4*)b, n4);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1;
  if (hipEventCreate(&e1) != hipSuccess) { hipEventDestroy(e0); return -1; }
  hipEventRecord(e0, st);
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy_stream, dim3(grid), dim3(256), 0, st, (const float4*)a, (float4*)b, n4);
  hipEventRecord(e1, st);
  hipError_t e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
  hipEventDestroy(e0); hipEventDestroy(e1);
  return e == hipSuccess ? 0 : -1;
}
void wv_launch_timed(const Ctx& c, const MGDev& m, int kernel, hipStream_t st) {
  if (kernel == 4) {   // Chebyshev sweep 2 on level 1 (the largest coarse level), gates open
    const int l = m.L > 1 ? 1 : 0;
    launch_coarse<0, false>(c, m, l, 1, 2, 1, st);
    return;
  }
  if (kernel == 3) { hipLaunchKernelGGL(k_copy_timed, dim3(nblocks(c.N, 256)), dim3(256), 0, st, c.work + 3 * c.N, m.ap, c.N); return; }
  if (kernel == 0) hipLaunchKernelGGL(k_mg_cheb_timed, dim3(nblocks(m.lv[0].count, 256)), dim3(256), 0, st, c, m);
  else if (kernel == 2) hipLaunchKernelGGL(k_mg_residual_timed, dim3(nblocks(m.lv[0].count, 256)), dim3(256), 0, st, c, m);
  else wv_launch_spmv_timed(c, st);
}

// ---- mixed-precision V-cycle -------------------------------------------------
__global__ void __launch_bounds__(256) k_mg_to_bf16(const float* __restrict__ src, uint16_t* __restrict__ dst, int64_t n) {
  int64_t i = GID(); if (i < n) dst[i] = f32_to_bf16(src[i]);
}
void wv_mg_to_bf16(const Ctx& s, const MGDev& m, hipStream_t st) {
  int64_t nr = 0; for (int l = 1; l < m.L; l++) nr += (int64_t)m.lv[l].count * m.lv[l].width;
  if (nr) hipLaunchKernelGGL(k_mg_to_bf16, dim3(nblocks(nr, 256)), dim3(256), 0, st, (const float*)m.rows, m.rows16, nr);
  int64_t ns = 7 * (int64_t)m.lv[0].count;
  hipLaunchKernelGGL(k_mg_to_bf16, dim3(nblocks(ns, 256)), dim3(256), 0, st, (const float*)m.snapshot, m.snap16, ns);
}
void wv_launch_mg_out64(const Ctx& s, const MGDev& m, int field, int phase, hipStream_t st);
void wv_mg_cycle_bf16(const Ctx& s, const MGDev& m, int phase, int smoothing, bool load, bool fp64_out, hipStream_t st) {
  if (load) hipLaunchKernelGGL(k_mg_load, dim3(nblocks(m.lv[0].count, 256)), dim3(256), 0, st, s, m, phase, (int)first_sweep_fused(m, 0));
  CycleBuilder<true> b{s, m, phase, smoothing, st};
  int final_field = b.cycle(0);
  if (fp64_out) wv_launch_mg_out64(s, m, final_field, phase, st);
  else {
    hipLaunchKernelGGL(k_mg_out<false>, dim3(nblocks(s.N, WV_RED)), dim3(WV_RED), 0, st, s, m, final_field, phase, 0);
    hipLaunchKernelGGL(k_mg_out<true>, dim3(1), dim3(WV_RED), 0, st, s, m, final_field, phase, 1);
  }
}
