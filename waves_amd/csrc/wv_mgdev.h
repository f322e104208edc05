// Multigrid device helpers shared by the V-cycle, the build and the
// standalone pressure solve (pressure_multigrid.glslinc:29-120).
#pragma once
#include "wv_kernels.h"

#define GID() ((int64_t)blockIdx.x * blockDim.x + threadIdx.x)

__device__ __forceinline__ int pmod(int x, int n) {
  if ((unsigned)x < (unsigned)n) return x;   // in range: the modulo is the identity
  const int y = x < 0 ? x + n : x - n;       // one period away: one add (almost every wrap of a stencil or transfer offset)
  if ((unsigned)y < (unsigned)n) return y;
  return x >= 0 ? x % n : (n - 1) - ((-1 - x) % n);
}
__device__ __forceinline__ int64_t at(const MGDev& m, int l, int field, int64_t i) { return m.lv[l].vec_off + (int64_t)field * m.lv[l].count + i; }
__device__ __forceinline__ I3 mdims(const MGDev& m, int l) { return mk(m.lv[l].n[0], m.lv[l].n[1], m.lv[l].n[2]); }
// pressure_multigrid.glslinc:36-47
__device__ __forceinline__ int mwrap(const MGDev& m, int l, I3 q) {
  I3 n = mdims(m, l); q.z = pmod(q.z, n.z);
  if (m.perx) q.x = pmod(q.x, n.x);
  if (q.x < 0 || q.x >= n.x || q.y < 0 || q.y >= n.y) return -1;
  return (int)index3(q, n);
}
__device__ __forceinline__ I3 mdelta(const MGDev& m, int l, int i, int j) {
  I3 n = mdims(m, l), a = coord(i, n), b = coord(j, n);
  I3 d = mk(b.x - a.x, b.y - a.y, b.z - a.z);
  if (m.perx) d.x = pmod(d.x + n.x / 2, n.x) - n.x / 2;
  d.z = pmod(d.z + n.z / 2, n.z) - n.z / 2;
  return d;
}
__device__ __forceinline__ int slot125(I3 d) { return d.x + 2 + 5 * (d.y + 2 + 5 * (d.z + 2)); }
__device__ __forceinline__ bool in_box2(I3 d) { return abs(d.x) <= 2 && abs(d.y) <= 2 && abs(d.z) <= 2; }
__device__ __forceinline__ I3 moffset(const MGDev& m, int l, int k) { const int* o = m.offsets + 3 * (m.lv[l].off_start + k); return mk(o[0], o[1], o[2]); }
__device__ __forceinline__ bool mactive(const MGDev& m, int l, int i) { return i >= 0 && m.vec[at(m, l, 3, i)] > .5f; }
__device__ __forceinline__ float snap(const MGDev& m, int field, int64_t i) { return m.snapshot[(int64_t)(field - 6) * m.lv[0].count + i]; }

// pressure_multigrid.glslinc:90-105
__device__ __forceinline__ float mapply(const MGDev& m, int l, int field, int i) {
  if (!mactive(m, l, i)) return m.vec[at(m, l, field, i)];
  float sum = 0; I3 q = coord(i, mdims(m, l));
  if (l == 0) {
    float here = m.vec[at(m, 0, field, i)];
#pragma unroll
    for (int f6 = 0; f6 < 6; f6++) {
      int axis = f6 >> 1, side = f6 & 1; I3 nb = q; addc(nb, axis, 2 * side - 1);
      int j = mwrap(m, 0, nb);
      float other = mactive(m, 0, j) ? m.vec[at(m, 0, field, j)] : 0;
      sum += snap(m, 6 + f6, i) * (here - other);
    }
    return sum;
  }
  int width = m.lv[l].width, N = m.lv[l].count; const float* row = m.rows + m.lv[l].row_off + i;
  for (int k = 0; k < width; k++) {
    I3 off = moffset(m, l, k); int j = mwrap(m, l, mk(q.x + off.x, q.y + off.y, q.z + off.z));
    sum += row[(int64_t)k * N] * m.vec[at(m, l, field, j >= 0 ? j : 0)];
  }
  return sum;
}
// The transfers' lattice position (float(j) + .5) / float(f) - .5 of fine
// coordinate j (pressure_multigrid.glslinc:108-120) for a coarsening factor f
// of 1 or 2: a division by a power of two is exact, so it is the product with
// 1 / f (the same value, without the division's instruction sequence)
__device__ __forceinline__ float lattice_pos(int j, int f) {
  const float x = (float)j + .5f;
  return (f == 2 ? x * .5f : (f == 1 ? x : x / (float)f)) - .5f;
}
// q / f for a coordinate q >= 0 and a factor f of 1 or 2 (a shift), else the division
__device__ __forceinline__ int fdiv(int q, int f) { return f == 2 ? q >> 1 : (f == 1 ? q : q / f); }
// pressure_multigrid.glslinc:108-120: transfers of fine node i (level l) to level l+1
__device__ __forceinline__ int mtransfers(const MGDev& m, int l, int i, int* ids, float* w) {
  const MGLevel& L = m.lv[l]; I3 n = mdims(m, l + 1), q = coord(i, mdims(m, l));
  if (L.aggregate) {
    int j = (int)index3(mk(fdiv(q.x, L.f[0]), fdiv(q.y, L.f[1]), fdiv(q.z, L.f[2])), n);
    if (!mactive(m, l + 1, j)) return 0; ids[0] = j; w[0] = 1; return 1;
  }
  int base[3]; float t[3]; int qq[3] = {q.x, q.y, q.z};
  for (int k = 0; k < 3; k++) { float pos = lattice_pos(qq[k], L.f[k]); base[k] = (int)floorf(pos); t[k] = fract(pos); }
  int size = 0;
  for (int z = 0; z < L.f[2]; z++) for (int y = 0; y < L.f[1]; y++) for (int x = 0; x < L.f[0]; x++) {
    int bit[3] = {x, y, z}; I3 nb = mk(base[0] + x, base[1] + y, base[2] + z); float ww = 1;
    for (int axis = 0; axis < 3; axis++) if (L.f[axis] > 1) ww *= bit[axis] == 1 ? t[axis] : 1 - t[axis];
    if (nb.y >= n.y) continue; if (nb.y < 0) nb.y = 0;
    nb.x = m.perx ? pmod(nb.x, n.x) : iclamp(nb.x, 0, n.x - 1); nb.z = pmod(nb.z, n.z);
    int j = (int)index3(nb, n); if (!mactive(m, l + 1, j) || ww == 0) continue;
    ids[size] = j; w[size] = ww; size++;
  }
  return size;
}
// 4th-kind Chebyshev sweep k, l1-Jacobi scaled (pressure_multigrid.glslinc:192-196):
// the new direction from the previous one and the residual r = b - A x; the
// sweep then writes x + d
__device__ __forceinline__ float cheb_dir(int k, float dprev, float dinv, float r) {
  return (k == 1 ? 0 : (float)(2 * k - 3) / (float)(2 * k + 1) * dprev) + (float)(8 * k - 4) / (float)(2 * k + 1) * dinv * r;
}
// Sweep 1 of a level whose x was just set to +0 by the kernel that also sets
// its b (the V-cycle's load, the PCG update, a restriction): A x is then +0 in
// every row (finite coefficients times +0, summed from +0; an inactive row
// takes x itself), so r = b - A x = b and the sweep needs no neighbour.  The
// producing kernel writes the sweep's d and x' = x + d = +0 + d itself, the
// same values as the sweep's launch.
__device__ __forceinline__ void cheb_first(float* d_out, float* x_out, float dinv, float b) {
  const float d = cheb_dir(1, 0.0f, dinv, b - 0.0f);
  *d_out = d; *x_out = 0.0f + d;
}
__device__ __forceinline__ bool mg_enabled(const Ctx& s, int phase) {
  const float* c = s.c;
  if (c[7] > .5f) return false;
  if (phase == 0) return (c[5] > .5f || c[57] == 2) && c[6] < .5f;
  return c[4] > .5f && c[6] < .5f;
}


// ---- BF16 operator storage (mixed-precision V-cycle) ------------------------
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// level-0 snapshot coefficient `field` (6..12) of cell i at the chosen precision
template <bool H>
__device__ __forceinline__ float snapc(const MGDev& m, int field, int64_t i) {
  if (H) return bf16_to_f32(m.snap16[(int64_t)(field - 6) * m.lv[0].count + i]);
  return snap(m, field, i);
}
// pressure_multigrid.glslinc:90-99 at level 0, coefficients at the chosen precision
template <bool H>
__device__ __forceinline__ float mapply0(const MGDev& m, int field, int i) {
  if (!mactive(m, 0, i)) return m.vec[at(m, 0, field, i)];
  float sum = 0; I3 q = coord(i, mdims(m, 0));
  float here = m.vec[at(m, 0, field, i)];
  // the six neighbours' flags and values loaded together (index 0 for a
  // neighbour outside the grid); mactive(j) selects the value
  float flag[6], value[6]; bool in[6];
#pragma unroll
  for (int f6 = 0; f6 < 6; f6++) {
    int axis = f6 >> 1, side = f6 & 1; I3 nb = q; addc(nb, axis, 2 * side - 1);
    const int j = mwrap(m, 0, nb); in[f6] = j >= 0;
    const int jj = j >= 0 ? j : 0;
    flag[f6] = m.vec[at(m, 0, 3, jj)]; value[f6] = m.vec[at(m, 0, field, jj)];
  }
#pragma unroll
  for (int f6 = 0; f6 < 6; f6++) {
    float other = in[f6] && flag[f6] > .5f ? value[f6] : 0;
    sum += snapc<H>(m, 6 + f6, i) * (here - other);
  }
  return sum;
}
