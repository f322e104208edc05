// PLIC reconstruction, interface geometry and coastal target on the device.
// Restates water/shaders/compute/plic_math.glslinc, pressure_geometry.glslinc,
// height_normal.glslinc and coastal_target.glslinc (same operation order).
#pragma once
#include "wv_device.h"

// plic_math.glslinc:3-23: unit-cube volume under n.x <= alpha
__device__ __forceinline__ float cube_volume(F3 n, float alpha) {
  alpha -= gmin(n.x, 0.0f) + gmin(n.y, 0.0f) + gmin(n.z, 0.0f);
  float x = fabsf(n.x), y = fabsf(n.y), z = fabsf(n.z);
  float total = x + y + z;
  if (alpha <= 0.0f) return 0.0f;
  if (alpha >= total) return 1.0f;
  float t;
  if (x > y) { t = x; x = y; y = t; }
  if (y > z) { t = y; y = z; z = t; }
  if (x > y) { t = x; x = y; y = t; }
  bool complement = alpha > 0.5f * total;
  t = complement ? total - alpha : alpha;
  float result;
  if (x < 1e-6f * total) {
    t -= x * 0.5f;
    if (y < 1e-6f * total) result = (t - y * 0.5f) / z;
    else if (t < y) result = t * t / (2 * y * z);
    else result = (t - 0.5f * y) / z;
  } else if (t < x) result = t * t * t / (6 * x * y * z);
  else if (t < y) result = (3 * t * t - 3 * t * x + x * x) / (6 * y * z);
  else if (x + y < z && t >= x + y) result = (t - 0.5f * (x + y)) / z;
  else {
    float ty = gmax(t - y, 0.0f), tz = gmax(t - z, 0.0f);
    // GLSL pow(v, 3.0) carries only exp2/log2 precision; the cube is the
    // product (v * v) * v, as in the oracle, and costs no libm call per step
    float numerator = x * (3 * t * t - 3 * t * x + x * x) - ty * ty * ty - tz * tz * tz;
    result = numerator / (6 * x * y * z);
  }
  return complement ? 1.0f - result : result;
}
// plic_math.glslinc:24-37: 25-step bisection for the plane constant
__device__ __forceinline__ float inverse_volume(F3 n, float fraction) {
  float lo = gmin(n.x, 0.0f) + gmin(n.y, 0.0f) + gmin(n.z, 0.0f);
  float hi = gmax(n.x, 0.0f) + gmax(n.y, 0.0f) + gmax(n.z, 0.0f);
  for (int it = 0; it < 25; it++) { float mid = 0.5f * (lo + hi); if (cube_volume(n, mid) < fraction) lo = mid; else hi = mid; }
  float alpha = 0.5f * (lo + hi), ctr = dot3(n, mkf(.5f, .5f, .5f));
  if (fraction == .5f) return ctr;
  if (fraction < .5f) return gmin(alpha, ctr);
  if (alpha <= ctr) alpha = ctr + 5.96046448e-8f * dot3(mkf(fabsf(n.x), fabsf(n.y), fabsf(n.z)), mkf(1, 1, 1));
  return alpha;
}

// inverse_volume on the eight lanes of a group (lanes 8g .. 8g + 7 of a wave,
// sub = lane & 7), all holding the same n and fraction.  Each round evaluates
// the midpoints of the next three bisection levels at once -- lane r takes
// node r of the round's tree, its interval halved from the round's (lo, hi)
// along r's path exactly as the serial loop would -- then every lane follows
// the decisions down the tree.  The midpoints, the tests and the order of the
// halvings are the serial loop's: the result is bit-identical, in 9 rounds of
// dependent work instead of 25.
__device__ __forceinline__ float inverse_volume8(F3 n, float fraction, int sub) {
  float lo = gmin(n.x, 0.0f) + gmin(n.y, 0.0f) + gmin(n.z, 0.0f);
  float hi = gmax(n.x, 0.0f) + gmax(n.y, 0.0f) + gmax(n.z, 0.0f);
  for (int done = 0; done < 25;) {
    const int levels = 25 - done >= 3 ? 3 : 25 - done;
    int test = 0;
    if (sub >= 1 && sub < (1 << levels)) {
      float l = lo, h = hi;
      const int depth = 31 - __clz(sub);
      for (int b = depth - 1; b >= 0; b--) { const float m = 0.5f * (l + h); if ((sub >> b) & 1) l = m; else h = m; }
      test = cube_volume(n, 0.5f * (l + h)) < fraction;
    }
    int node = 1;
    for (int lev = 0; lev < levels; lev++) {
      const int t = __shfl(test, node, 8);
      const float mid = 0.5f * (lo + hi);
      if (t) lo = mid; else hi = mid;
      node = 2 * node + t;
    }
    done += levels;
  }
  float alpha = 0.5f * (lo + hi), ctr = dot3(n, mkf(.5f, .5f, .5f));
  if (fraction == .5f) return ctr;
  if (fraction < .5f) return gmin(alpha, ctr);
  if (alpha <= ctr) alpha = ctr + 5.96046448e-8f * dot3(mkf(fabsf(n.x), fabsf(n.y), fabsf(n.z)), mkf(1, 1, 1));
  return alpha;
}

// ---- coastal_target.glslinc ----
__device__ __forceinline__ float smooth_ramp(float q) { q = gclamp(q, 0, 1); return q * q * q * (10 + q * (-15 + 6 * q)); }
__device__ __forceinline__ float ramped_amplitude(const Ctx& s, float time) {
  return s.wave[0] * smooth_ramp(time * s.wave[2] / (6.28318530718f * s.forc[3]));
}
__device__ __forceinline__ void coastal_surface(const Ctx& s, float x, float time, float& eta_out, float& slope_out) {
  float a = ramped_amplitude(s, time), k = s.wave[1], w = s.wave[2], h = s.wave[3];
  float sigma = tanhf(k * h), second = k * a * a * (3 - sigma * sigma) / (4 * sigma * sigma * sigma);
  if (s.stokes2 < .5f) second = 0;
  float phase = k * (x - s.forc[0]) - w * time;
  eta_out = a * cosf(phase) + second * cosf(2 * phase);
  slope_out = -k * a * sinf(phase) - 2 * k * second * sinf(2 * phase);
}
__device__ __forceinline__ F3 coastal_velocity(const Ctx& s, F3 world, float time) {
  float a = ramped_amplitude(s, time), k = s.wave[1], w = s.wave[2], h = s.wave[3];
  float sh = sinhf(k * h), phase = k * (world.x - s.forc[0]) - w * time, y = world.y + h;
  float first = a * w / sh, second = s.stokes2 > .5f ? .75f * a * a * w * k / powf(sh, 4.0f) : 0;
  return mkf(first * coshf(k * y) * cosf(phase) + second * coshf(2 * k * y) * cosf(2 * phase),
             first * sinhf(k * y) * sinf(phase) + second * sinhf(2 * k * y) * sinf(2 * phase), 0);
}
// One component of coastal_velocity, the same expression evaluated for that
// component only
__device__ __forceinline__ float coastal_velocity_axis(const Ctx& s, F3 world, float time, int axis) {
  if (axis == 2) return 0.0f;
  float a = ramped_amplitude(s, time), k = s.wave[1], w = s.wave[2], h = s.wave[3];
  float sh = sinhf(k * h), phase = k * (world.x - s.forc[0]) - w * time, y = world.y + h;
  float first = a * w / sh, second = s.stokes2 > .5f ? .75f * a * a * w * k / powf(sh, 4.0f) : 0;
  if (axis == 0) return first * coshf(k * y) * cosf(phase) + second * coshf(2 * k * y) * cosf(2 * phase);
  return first * sinhf(k * y) * sinf(phase) + second * sinhf(2 * k * y) * sinf(2 * phase);
}
__device__ __forceinline__ F3 coastal_face_position(const Ctx& s, I3 q, int axis) { F3 p = center(s, q); addc(p, axis, -(.5f * s.h[axis])); return p; }
__device__ __forceinline__ float4 coastal_plane(const Ctx& s, I3 q, float time) {
  float se, ss; coastal_surface(s, center(s, q).x, time, se, ss);
  F3 normal = mkf(-ss * s.h[0], s.h[1], 0);
  return make_float4(normal.x, normal.y, normal.z, dot3(normal, mkf(.5f, .5f, .5f)) + se - center(s, q).y);
}
__device__ __forceinline__ float coastal_fraction(const Ctx& s, I3 q, float time) {
  float4 eq = coastal_plane(s, q, time);
  return gclamp(cube_volume(mkf(eq.x, eq.y, eq.z), eq.w), 0, 1);
}
__device__ __forceinline__ float coastal_rate(const Ctx& s, float x) {
  float left = 1 - smooth_ramp((x - s.forc[0]) / s.forc[1]);
  float right = s.shape == 6 ? smooth_ramp((x - (s.o[0] + (float)s.nx * s.h[0] - s.forc[1])) / s.forc[1]) : 0;
  return s.forc[2] * gmax(left, s.right_abs * right);
}

// pressure_geometry.glslinc:39-47 (fractions of the geometry bank).  The
// *_known forms take solid(s, q), which a caller that has it already passes
// in (solid() wraps q itself, so the flag of q and of its wrapped cell agree).
__device__ __forceinline__ float geometry_fraction_known(const Ctx& s, I3 q, int gbank, bool sq) {
  q = wrap_cell(s, q);
  if (q.x < 0 && !sq) return coastal_fraction(s, q, s.c[1] + (gbank == candidate(s) ? timestep(s) : 0));
  if (sq) { q.x = iclamp(q.x, 0, s.nx - 1); q.y = iclamp(q.y, 0, s.ny - 1); q.z = iclamp(q.z, 0, s.nz - 1); }
  if (outside(s, q)) return 0.0f;
  return getv(s, gbank, index3(q, dims(s)));
}
__device__ __forceinline__ float geometry_fraction(const Ctx& s, I3 q, int gbank) { return geometry_fraction_known(s, q, gbank, solid(s, q)); }
// coupled_motion.glslinc:36-44 (fractions of the current transport sweep)
__device__ __forceinline__ float transport_f_known(const Ctx& s, I3 q, bool sq) {
  q = wrap_cell(s, q);
  if (q.x < 0 && !sq) return coastal_fraction(s, q, s.c[1]);
  if (outside(s, q) || sq) return 0;
  int64_t i = index3(q, dims(s)); int sweep = (int)s.c[21];
  return sweep == 0 ? getv(s, accepted(s), i) : (sweep == 1 ? getv(s, candidate(s), i) : s.motion[i]);
}
__device__ __forceinline__ float transport_f(const Ctx& s, I3 q) { return transport_f_known(s, q, solid(s, q)); }
// source 0: geometry bank; source 1: transport sweep
__device__ __forceinline__ float hf_fraction(const Ctx& s, I3 q, int source, int gbank) {
  return source == 1 ? transport_f(s, q) : geometry_fraction(s, q, gbank);
}
__device__ __forceinline__ float hf_fraction_known(const Ctx& s, I3 q, int source, int gbank, bool sq) {
  return source == 1 ? transport_f_known(s, q, sq) : geometry_fraction_known(s, q, gbank, sq);
}

// pressure_geometry.glslinc:72-90: weighted Youngs gradient, one-sided next to solids
// Every cell the stencil touches lies in the 3x3x3 block around q: its solid
// flags and fractions are read first, 27 independent loads issued together,
// and the loops below (fully unrolled, constant block indices) run on
// registers.  The values and the order of the arithmetic are the reference's.
// the gradient from the block's solid flags sb and fractions fb (cell c at
// offset (c % 3 - 1, c / 3 % 3 - 1, c / 9 - 1))
__device__ __forceinline__ F3 youngs_from_block(const Ctx& s, const bool* sb, const float* fb) {
  F3 g = mkf(0, 0, 0);
#pragma unroll
  for (int axis = 0; axis < 3; axis++) {
    const int b = (axis + 1) % 3, d = (axis + 2) % 3;
    int weights = 0;
    float acc = 0;
#pragma unroll
    for (int j = -1; j <= 1; j++)
#pragma unroll
      for (int k = -1; k <= 1; k++) {
        int o[3] = {0, 0, 0}; o[b] = j; o[d] = k;
        const int cm = (o[0] + 1) + 3 * ((o[1] + 1) + 3 * (o[2] + 1));
        const int step = axis == 0 ? 1 : (axis == 1 ? 3 : 9);
        const int cl = cm - step, cr = cm + step;
        if (sb[cm]) continue;
        bool left = !sb[cl], right = !sb[cr]; if (!left && !right) continue;
        float difference;
        if (left && right) difference = fb[cl] - fb[cr];
        else if (left) difference = 2 * (fb[cl] - fb[cm]);
        else difference = 2 * (fb[cm] - fb[cr]);
        int weight = (j == 0 ? 2 : 1) * (k == 0 ? 2 : 1); weights += weight;
        acc += (float)weight * difference / (32 * s.h[axis]);
      }
    if (weights > 0) acc *= 16.0f / (float)weights;
    setc(g, axis, acc);
  }
  return g;
}
__device__ __forceinline__ F3 interface_gradient(const Ctx& s, I3 q, int source, int gbank) {
  bool sb[27]; float fb[27];
#pragma unroll
  for (int c = 0; c < 27; c++) {
    I3 p = mk(q.x + c % 3 - 1, q.y + (c / 3) % 3 - 1, q.z + c / 9 - 1);
    sb[c] = solid(s, p); fb[c] = hf_fraction_known(s, p, source, gbank, sb[c]);
  }
  return youngs_from_block(s, sb, fb);
}
// The same gradient on the eight lanes of a group (lanes 8g .. 8g + 7 of a
// wave, sub = lane & 7, all for the same cell): lane sub evaluates the block's
// cells sub, sub + 8, ..., which every lane then gathers by shuffles.
__device__ __forceinline__ F3 interface_gradient8(const Ctx& s, I3 q, int source, int gbank, int sub) {
  int msb[4]; float mfb[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int c = sub + 8 * u < 27 ? sub + 8 * u : 26;
    const I3 p = mk(q.x + c % 3 - 1, q.y + (c / 3) % 3 - 1, q.z + c / 9 - 1);
    const bool sc = solid(s, p); msb[u] = sc; mfb[u] = hf_fraction_known(s, p, source, gbank, sc);
  }
  bool sb[27]; float fb[27];
#pragma unroll
  for (int c = 0; c < 27; c++) { sb[c] = __shfl(msb[c >> 3], c & 7, 8) != 0; fb[c] = __shfl(mfb[c >> 3], c & 7, 8); }
  return youngs_from_block(s, sb, fb);
}
__device__ __forceinline__ bool hf_valid(const Ctx& s, I3 q) { q = wrap_cell(s, q); return !outside(s, q) && !solid(s, q); }
// height_normal.glslinc:6-34 (numerics flag bit 0)
__device__ __forceinline__ F3 height_normal(const Ctx& s, I3 q, F3 youngs, int source, int gbank) {
  F3 scaled = mkf(fabsf(youngs.x) * s.h[0], fabsf(youngs.y) * s.h[1], fabsf(youngs.z) * s.h[2]);
  int axis = scaled.x > scaled.y ? 0 : 1; if (scaled.z > compc(scaled, axis)) axis = 2;
  if (compc(scaled, axis) < 1e-12f) return youngs;
  int b = (axis + 1) % 3, d = (axis + 2) % 3; float orientation = gsign(compc(youngs, axis));
  float heights[5];
  for (int column = 0; column < 5; column++) {
    I3 base = q;
    if (column == 1) addc(base, b, -1); if (column == 2) addc(base, b, 1);
    if (column == 3) addc(base, d, -1); if (column == 4) addc(base, d, 1);
    float sum = 0, prev = orientation > 0 ? 1 : 0;
    for (int t = -4; t <= 4; t++) {
      I3 cell = base; addc(cell, axis, t); if (!hf_valid(s, cell)) return youngs;
      float f = hf_fraction(s, cell, source, gbank);
      if (isnan(f) || isinf(f) || f < 0 || f > 1) return youngs;
      if (orientation * (f - prev) > 2e-6f) return youngs;
      if (t == -4 && f != (orientation > 0 ? 1 : 0)) return youngs;
      if (t == 4 && f != (orientation > 0 ? 0 : 1)) return youngs;
      sum += f; prev = f;
    }
    heights[column] = sum * s.h[axis];
  }
  F3 r = mkf(0, 0, 0); setc(r, axis, orientation);
  setc(r, b, -(heights[2] - heights[1]) / (2 * s.h[b]));
  setc(r, d, -(heights[4] - heights[3]) / (2 * s.h[d]));
  return r;
}
// pressure_geometry.glslinc:91-98 / coupled_motion.glslinc:50-56
__device__ __forceinline__ F3 normal_from_gradient(const Ctx& s, I3 q, F3 g, int source, int gbank) {
  if (s.flags & 1) g = height_normal(s, q, g, source, gbank);
  if (len3(g) < 1e-9f) g = mkf(0, 1, 0);
  float L = len3(g);
  F3 n = mkf(g.x / L * s.h[0], g.y / L * s.h[1], g.z / L * s.h[2]);
  float l1 = dot3(mkf(fabsf(n.x), fabsf(n.y), fabsf(n.z)), mkf(1, 1, 1));
  n.x /= l1; n.y /= l1; n.z /= l1;
  return n;
}
__device__ __forceinline__ float4 plane_from_gradient(const Ctx& s, I3 q, F3 g, int source, int gbank, float f) {
  F3 n = normal_from_gradient(s, q, g, source, gbank);
  return make_float4(n.x, n.y, n.z, inverse_volume(n, f));
}
// the plane of an interface cell on the eight lanes of a group
__device__ __forceinline__ float4 plane8(const Ctx& s, I3 q, int source, int gbank, float f, int sub) {
  F3 n = normal_from_gradient(s, q, interface_gradient8(s, q, source, gbank, sub), source, gbank);
  return make_float4(n.x, n.y, n.z, inverse_volume8(n, f, sub));
}

// pressure_geometry.glslinc:99-101
__device__ __forceinline__ float segment_distance(F3 p, F3 a, F3 b) {
  F3 ab = mkf(b.x - a.x, b.y - a.y, b.z - a.z), pa = mkf(p.x - a.x, p.y - a.y, p.z - a.z);
  float t = gclamp(dot3(pa, ab) / gmax(dot3(ab, ab), 1e-30f), 0, 1);
  return len3(mkf(p.x - (a.x + t * ab.x), p.y - (a.y + t * ab.y), p.z - (a.z + t * ab.z)));
}
// Vertex lists of the polygon routines (at most 12 vertices), one list per
// thread in LDS: coordinate c of vertex j at p[(3 j + c) * STRIDE] with p the
// thread's slot, so consecutive threads use consecutive banks.  As private
// arrays indexed by a running count they would live in scratch memory.
template <int STRIDE>
struct Verts {
  float* p;
  __device__ __forceinline__ F3 get(int j) const { return mkf(p[(3 * j) * STRIDE], p[(3 * j + 1) * STRIDE], p[(3 * j + 2) * STRIDE]); }
  __device__ __forceinline__ void put(int j, F3 a) const { p[(3 * j) * STRIDE] = a.x; p[(3 * j + 1) * STRIDE] = a.y; p[(3 * j + 2) * STRIDE] = a.z; }
};
constexpr int VERT_FLOATS = 36;  // floats of one list per thread
// pressure_geometry.glslinc:105-120: distance to the PLIC polygon of a cell
template <int STRIDE>
__device__ __forceinline__ float polygon_distance(const Ctx& s, F3 point, float4 eq, const Verts<STRIDE>& vert) {
  F3 h = mkf(s.h[0], s.h[1], s.h[2]);
  F3 normal = mkf(eq.x / h.x, eq.y / h.y, eq.z / h.z);
  float norm2 = dot3(normal, normal), step = dot3(normal, point) - eq.w;
  float signed_d = (dot3(normal, point) - eq.w) / sqrtf(norm2);
  F3 pr = mkf(point.x - normal.x * step / norm2, point.y - normal.y * step / norm2, point.z - normal.z * step / norm2);
  if (pr.x >= 0 && pr.y >= 0 && pr.z >= 0 && pr.x <= h.x && pr.y <= h.y && pr.z <= h.z) return fabsf(signed_d);
  int count = 0;
  float e3[3] = {eq.x, eq.y, eq.z};
  for (int axis = 0; axis < 3; axis++) for (int j = 0; j < 2; j++) for (int k = 0; k < 2; k++) {
    F3 a = mkf(0, 0, 0); setc(a, (axis + 1) % 3, (float)j); setc(a, (axis + 2) % 3, (float)k);
    float den = e3[axis]; if (fabsf(den) < 1e-15f) continue;
    float t = (eq.w - dot3(mkf(eq.x, eq.y, eq.z), a)) / den;
    if (t >= 0 && t <= 1) { setc(a, axis, t); vert.put(count++, mkf(a.x * h.x, a.y * h.y, a.z * h.z)); }
  }
  float best = 1e10f;
  for (int j = 0; j < count; j++) {
    const F3 vj = vert.get(j);
    best = gmin(best, len3(mkf(point.x - vj.x, point.y - vj.y, point.z - vj.z)));
    for (int k = 0; k < j; k++) best = gmin(best, segment_distance(point, vj, vert.get(k)));
  }
  return best;
}
// pressure_geometry.glslinc:124-132: the polygon in `poly` clipped by a plane
// into `out` (the reference clips in place through a copy: same vertices)
template <int STRIDE>
__device__ __forceinline__ int clip_polygon(const Verts<STRIDE>& poly, int count, float4 eq, const Verts<STRIDE>& out) {
  int size = 0; F3 n = mkf(eq.x, eq.y, eq.z);
  for (int j = 0; j < count; j++) {
    F3 a = poly.get(j), b = poly.get((j + 1) % count);
    float da = dot3(n, a) - eq.w, db = dot3(n, b) - eq.w;
    if (da <= 0 && size < 12) out.put(size++, a);
    if (((da < 0 && db > 0) || (da > 0 && db < 0)) && size < 12) {
      float t = da / (da - db);
      out.put(size++, mkf(a.x * (1 - t) + b.x * t, a.y * (1 - t) + b.y * t, a.z * (1 - t) + b.z * t));
    }
  }
  return size;
}
// pressure_geometry.glslinc:133-156: represented phase transitions on a cell face
template <int STRIDE>
__device__ __forceinline__ float transition_distance(const Ctx& s, F3 point, I3 left, int axis, int gbank, const Verts<STRIDE>& poly, const Verts<STRIDE>& spare) {
  I3 right = left; addc(right, axis, 1);
  I3 l = wrap_cell(s, left), r = wrap_cell(s, right);
  if (outside(s, l) || outside(s, r) || solid(s, l) || solid(s, r)) return 1e10f;
  float fl = geometry_fraction(s, l, gbank), fr = geometry_fraction(s, r, gbank);
  if ((fl <= 0 && fr <= 0) || (fl >= 1 && fr >= 1)) return 1e10f;
  float4 a = s.plane[index3(l, dims(s))], b = s.plane[index3(r, dims(s))];
  if (fl <= 0) a = make_float4(0, 0, 0, -1); if (fl >= 1) a = make_float4(0, 0, 0, 1);
  if (fr <= 0) b = make_float4(0, 0, 0, -1); if (fr >= 1) b = make_float4(0, 0, 0, 1);
  b.w += axis == 0 ? b.x : (axis == 1 ? b.y : b.z);
  float best = 1e10f; F3 h = mkf(s.h[0], s.h[1], s.h[2]);
  for (int side = 0; side < 2; side++) {
    int j = (axis + 1) % 3, k = (axis + 2) % 3;
    for (int v = 0; v < 4; v++) { F3 c = mkf(0, 0, 0); setc(c, axis, 1); setc(c, j, (v == 1 || v == 2) ? 1 : 0); setc(c, k, v >= 2 ? 1 : 0); poly.put(v, c); }
    float4 first = side == 0 ? a : b;
    float4 second = side == 0 ? make_float4(-b.x, -b.y, -b.z, -b.w) : make_float4(-a.x, -a.y, -a.z, -a.w);
    int count = clip_polygon(poly, 4, first, spare); count = clip_polygon(spare, count, second, poly);
    if (count == 0) continue;
    F3 projected = point; setc(projected, axis, compc(h, axis));
    F3 local = mkf(projected.x / h.x, projected.y / h.y, projected.z / h.z);
    if (local.x >= 0 && local.y >= 0 && local.z >= 0 && local.x <= 1 && local.y <= 1 && local.z <= 1 &&
        dot3(mkf(first.x, first.y, first.z), local) <= first.w && dot3(mkf(second.x, second.y, second.z), local) <= second.w)
      best = gmin(best, fabsf(compc(point, axis) - compc(h, axis)));
    for (int v = 0; v < count; v++) {
      F3 p0 = poly.get(v), p1 = poly.get((v + 1) % count);
      best = gmin(best, segment_distance(point, mkf(p0.x * h.x, p0.y * h.y, p0.z * h.z), mkf(p1.x * h.x, p1.y * h.y, p1.z * h.z)));
    }
  }
  return best;
}
// pressure_geometry.glslinc:157-182: exact signed distance at a cell centre
// The value of a cell that needs no search (an interface cell, or one whose
// six neighbours all lie on its own side), else `seed` is set.
__device__ __forceinline__ float distance_head(const Ctx& s, I3 q, int gbank, bool& seed) {
  float f = geometry_fraction(s, q, gbank); F3 h = mkf(s.h[0], s.h[1], s.h[2]);
  seed = false;
  if (f > 0 && f < 1) {
    float4 e = s.plane[index3(q, dims(s))];
    return (dot3(mkf(e.x, e.y, e.z), mkf(.5f, .5f, .5f)) - e.w) / len3(mkf(e.x / h.x, e.y / h.y, e.z / h.z));
  }
  bool needs_seed = false;
  for (int axis = 0; axis < 3; axis++) for (int sign = -1; sign <= 1; sign += 2) {
    I3 nb = q; addc(nb, axis, sign); nb = wrap_cell(s, nb); if (solid(s, nb) || outside(s, nb)) continue;
    float other = geometry_fraction(s, nb, gbank); bool inside = other >= 1;
    if (other > 0 && other < 1) { float4 e = s.plane[index3(nb, dims(s))]; inside = dot3(mkf(e.x, e.y, e.z), mkf(.5f, .5f, .5f)) < e.w; }
    if (inside != (f >= 1)) needs_seed = true;
  }
  seed = needs_seed;
  return (f >= 1 ? -1 : 1) * len3(h) * 4;
}
// Neighbour k (x-major over the 3x3x3 block, the search order) of a seeded
// cell: the distance to its box and the least of its polygon and transition
// distances; +inf where the neighbour is outside or solid.
template <int STRIDE>
__device__ __forceinline__ void distance_neighbour(const Ctx& s, I3 q, int k, int gbank, float& box, float& m, const Verts<STRIDE>& va, const Verts<STRIDE>& vb) {
  const float inf = __int_as_float(0x7f800000);
  box = inf; m = inf;
  I3 nbr = mk(q.x + k / 9 - 1, q.y + (k / 3) % 3 - 1, q.z + k % 3 - 1), wr = wrap_cell(s, nbr);
  if (outside(s, wr) || solid(s, wr)) return;
  F3 h = mkf(s.h[0], s.h[1], s.h[2]);
  float fraction = geometry_fraction(s, wr, gbank);
  F3 point = mkf(((float)(q.x - nbr.x) + .5f) * h.x, ((float)(q.y - nbr.y) + .5f) * h.y, ((float)(q.z - nbr.z) + .5f) * h.z);
  F3 d = mkf(point.x - gclamp(point.x, 0, h.x), point.y - gclamp(point.y, 0, h.y), point.z - gclamp(point.z, 0, h.z));
  box = len3(d);
  float c = inf;
  if (fraction > 0 && fraction < 1) c = gmin(c, polygon_distance(s, point, s.plane[index3(wr, dims(s))], va));
  for (int axis = 0; axis < 3; axis++) c = gmin(c, transition_distance(s, point, nbr, axis, gbank, va, vb));
  m = c;
}
// pressure_geometry.glslinc:157-182.  The search keeps a running best and
// skips a neighbour whose box is no nearer; min(best, m_k) equals the
// reference's chain of mins over the neighbour's candidates (none is NaN-
// sticky: gmin keeps the left operand when the right one is NaN).
__device__ __forceinline__ float distance_step(float best, float box, float m) { return box >= best ? best : gmin(best, m); }
