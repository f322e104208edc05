// Face-velocity sampling, MacCormack traces, CFL rates and air extension on the
// device.  Restates water/shaders/compute/coupled_motion.glslinc:15-230.
#pragma once
#include "wv_geom.h"

__device__ __forceinline__ F3 periodic_delta(const Ctx& s, I3 a, I3 b) {
  F3 d = mkf((float)(a.x - b.x), (float)(a.y - b.y), (float)(a.z - b.z));
  d.z -= roundf(d.z / (float)s.nz) * (float)s.nz;
  if (s.perx > .5f) d.x -= roundf(d.x / (float)s.nx) * (float)s.nx;
  return mkf(d.x * s.h[0], d.y * s.h[1], d.z * s.h[2]);
}
__device__ __forceinline__ bool wall_face(const Ctx& s, I3 q, int axis) { I3 l = q; addc(l, axis, -1); return solid(s, q) || solid(s, l); }
__device__ __forceinline__ bool face_in_bounds(const Ctx& s, I3 q, int axis) {
  I3 n = dims(s); addc(n, axis, 1);
  return q.x >= 0 && q.y >= 0 && q.z >= 0 && q.x < n.x && q.y < n.y && q.z < n.z;
}
__device__ __forceinline__ I3 wrap_face(const Ctx& s, I3 q) { q.z = pidx(q.z, s.nz); if (s.perx > .5f) q.x = pidx(q.x, s.nx); return q; }
// coupled_motion.glslinc:28-35: accepted face velocity with the ocean ghost
__device__ __forceinline__ float old_face(const Ctx& s, I3 q, int axis) {
  q = wrap_face(s, q);
  if (q.x < 0 && !wall_face(s, q, axis)) return coastal_velocity_axis(s, coastal_face_position(s, q, axis), s.c[1], axis);
  if (!face_in_bounds(s, q, axis) || wall_face(s, q, axis)) return 0;
  return getv(s, accepted(s), face_index(s, q, axis));
}
__device__ __forceinline__ float face_courant(const Ctx& s, I3 q, int axis) { return old_face(s, q, axis) * timestep(s) / s.h[axis]; }

// coupled_motion.glslinc:66-85: nearest wall within three layers.  The search
// returns the chosen mirror as an offset code, (axis << 4) | (8 if the step is
// positive) | k with the mirror k faces along axis, or -1 when no wall is in
// reach.  It reads solid() only, and solid() does not depend on z (the bed is
// a function of x and y, and z wraps), so the code for a face depends on its
// component, x and y alone: Ctx::reflect holds it for every in-range face at
// z = 0, filled once when the solver is created (k_reflect_table).
__device__ __forceinline__ int reflect_search(const Ctx& s, I3 q, int component) {
  float best = 1e20f; int code = -1;
  for (int axis = 0; axis < 3; axis++) for (int sign = -1; sign <= 1; sign += 2) {
    for (int layer = 1; layer <= 3; layer++) {
      I3 edge = q; addc(edge, axis, sign * layer); edge = wrap_face(s, edge);
      if (!face_in_bounds(s, edge, component)) break;
      I3 left = edge; addc(left, component, -1);
      bool sl = solid(s, left), sr = solid(s, edge);
      bool boundary = axis == component ? (sl != sr) : (!sl && !sr);
      if (!boundary) continue;
      const int k = 2 * layer - (axis == component ? 0 : 1);
      I3 mirror = q; addc(mirror, axis, sign * k); mirror = wrap_face(s, mirror);
      float distance = ((float)layer - (axis == component ? 0 : .5f)) * s.h[axis];
      if (face_in_bounds(s, mirror, component) && !wall_face(s, mirror, component) && distance < best) {
        code = (axis << 4) | (sign > 0 ? 8 : 0) | k; best = distance;
      }
      break;
    }
  }
  return code;
}
// the mirror face and parity of a face between two solid cells (q wrapped, in range)
__device__ __forceinline__ bool reflect_voxel_donor(const Ctx& s, I3& q, int component, float& parity) {
  const int code = s.reflect[((int64_t)component * (s.ny + 1) + q.y) * (s.nx + 1) + q.x];
  parity = 1;
  if (code < 0) return false;
  const int axis = code >> 4;
  addc(q, axis, (code & 8) ? (code & 7) : -(code & 7)); q = wrap_face(s, q);
  parity = axis == component ? -1 : 1;
  return true;
}
// solid() for a cell whose z is already in [0, nz) without periodic x, where
// the periodic wrap is the identity
__device__ __forceinline__ bool solid_direct(const Ctx& s, I3 q) {
  if (q.y < 0 || q.x >= s.nx) return true;
  if (q.y >= s.ny) return false;
  return solid_formula(s, q);
}
// coupled_motion.glslinc:86-111.  The first branch is the common case of the
// general code below, taken without its reflections and periodic wraps: a face
// inside the grid (x >= 0, no periodic x, z and its left neighbour in range)
// between two non-solid cells reads the stored value with parity 1.
__device__ __forceinline__ float velocity_donor(const Ctx& s, I3 q, int axis, bool forward, bool& valid) {
  if (s.perx < .5f && q.x >= 0 && q.y >= 0 && q.z >= (axis == 2 ? 1 : 0) && q.z < s.nz &&
      q.x < s.nx + (axis == 0 ? 1 : 0) && q.y < s.ny + (axis == 1 ? 1 : 0)) {
    I3 l = q; addc(l, axis, -1);
    if (!solid_direct(s, l) && !solid_direct(s, q)) {
      int acc = accepted(s);
      int64_t f = face_local(s, q, axis); valid = s.motion[valid_base(s, acc) + f] > .5f;
      return forward ? s.motion[forward_base(s) + f] : getv(s, acc, U_(s) + f);
    }
  }
  float parity = 1;
  if (q.y < 0) { q.y = -q.y - (axis == 1 ? 0 : 1); if (axis == 1) parity = -parity; }
  if (s.perx < .5f) {
    if (q.x >= s.nx + (axis == 0 ? 1 : 0)) { q.x = 2 * s.nx - q.x - (axis == 0 ? 0 : 1); if (axis == 0) parity = -parity; }
  }
  q = wrap_face(s, q);
  if (q.x < 0 && q.y >= 0 && q.y < s.ny && !wall_face(s, q, axis)) {
    valid = true; return parity * coastal_velocity_axis(s, coastal_face_position(s, q, axis), s.c[1], axis);
  }
  if (!face_in_bounds(s, q, axis)) { valid = false; return 0; }
  I3 l = q; addc(l, axis, -1); bool left = solid(s, l), right = solid(s, q);
  if (left != right) { valid = true; return 0; }
  if (left && right) {
    float reflection;
    if (!reflect_voxel_donor(s, q, axis, reflection)) { valid = false; return 0; }
    parity *= reflection;
  }
  int acc = accepted(s);
  int64_t f = face_local(s, q, axis); valid = s.motion[valid_base(s, acc) + f] > .5f;
  return parity * (forward ? s.motion[forward_base(s) + f] : getv(s, acc, face_index(s, q, axis)));
}
// one lattice corner of a staggered sample: value and support of the face at lo + (xx, y, z)
__device__ __forceinline__ float sample_tap(const Ctx& s, I3 lo, int xx, int y, int z, int axis, bool forward, bool direct,
                                            I3 fn, int acc, bool& supported) {
  if (direct) {
    // a direct corner reads the stored face at the wrapped z (sample_at)
    int64_t f = face_local(s, mk(lo.x + xx, lo.y + y, pidx(lo.z + z, s.nz)), axis);
    supported = s.motion[valid_base(s, acc) + f] > .5f;
    return forward ? s.motion[forward_base(s) + f] : getv(s, acc, U_(s) + f);
  }
  return velocity_donor(s, mk(lo.x + xx, lo.y + y, lo.z + z), axis, forward, supported);
}
// coupled_motion.glslinc:112-137: staggered trilinear sample in fma difference
// form.  sample_at places the sample on the component's lattice and decides
// whether all eight corners are direct reads; sample_direct then issues the
// sixteen loads together, the general path visits the corners as four
// x-pairs in a loop that is not unrolled (one copy of the donor code).  The
// corner order, the support and limit updates and the interpolation are the
// reference's on both paths.
struct SampleAt { I3 lo; F3 t; bool direct; };
__device__ __forceinline__ SampleAt sample_at(const Ctx& s, F3 x, int axis) {
  SampleAt a;
  F3 lattice = mkf(x.x / s.h[0] - .5f, x.y / s.h[1] - .5f, x.z / s.h[2] - .5f);
  addc(lattice, axis, .5f);
  I3 lo = mk((int)floorf(lattice.x), (int)floorf(lattice.y), (int)floorf(lattice.z));
  a.lo = lo;
  a.t = mkf(fract(lattice.x), fract(lattice.y), fract(lattice.z));
  bool interior = (s.flags & 8) != 0 && lo.x >= 1 && lo.y >= 1 && lo.z >= 1 && lo.x + 1 < s.nx && lo.y + 1 < s.ny && lo.z + 1 < s.nz &&
                  !solid(s, mk(lo.x + 1, lo.y - 1, lo.z));
  // All eight corners read the stored face with parity 1 (in range in x and
  // y, no periodic x, both cells of every corner face not solid): the bed test
  // is monotone in y, so the lowest cell of each touched column decides for
  // the column.  Neither the bed nor solid() depends on z (solid() wraps it),
  // so a corner across the periodic z edge is the stored face at the wrapped
  // z, which is what velocity_donor's general path reads there (wrap_face,
  // then both cells non-solid).
  bool fluid = s.perx < .5f && lo.x >= 0 && lo.y >= 0 &&
               lo.x + 1 < s.nx + (axis == 0 ? 1 : 0) && lo.y + 1 < s.ny + (axis == 1 ? 1 : 0);
  if (fluid) {
    int ylow = lo.y - (axis == 1 ? 1 : 0);
    for (int cx = lo.x - (axis == 0 ? 1 : 0); cx <= lo.x + 1; cx++) fluid = fluid && !solid_direct(s, mk(cx, ylow, lo.z));
  }
  a.direct = interior || fluid;
  return a;
}
__device__ __forceinline__ void corner_update(F3 t, int c8, float v, bool sup, bool& valid, float& lim_lo, float& lim_hi) {
  int xx = c8 & 1, y = (c8 >> 1) & 1, z = c8 >> 2;
  float wx = (1 - t.x) * (1 - (float)xx) + t.x * (float)xx;
  float wy = (1 - t.y) * (1 - (float)y) + t.y * (float)y;
  float wz = (1 - t.z) * (1 - (float)z) + t.z * (float)z;
  if (wx * wy * wz > 0) { valid = valid && sup; lim_lo = gmin(lim_lo, v); lim_hi = gmax(lim_hi, v); }
}
__device__ __forceinline__ float sample_direct(const Ctx& s, const SampleAt& at, int axis, bool forward, bool& valid, float& lim_lo, float& lim_hi) {
  I3 fn = dims(s); addc(fn, axis, 1);
  int acc = accepted(s);
  // the two z layers of the corners, wrapped over the periodic z edge
  const int64_t fz0 = face_local(s, mk(at.lo.x, at.lo.y, pidx(at.lo.z, s.nz)), axis);
  const int64_t fz1 = face_local(s, mk(at.lo.x, at.lo.y, pidx(at.lo.z + 1, s.nz)), axis);
  int64_t vb = valid_base(s, acc), fb = forward_base(s), ub = U_(s);
  float v[8]; bool sup[8];
#pragma unroll
  for (int c8 = 0; c8 < 8; c8++) {
    int64_t f = (c8 >> 2 ? fz1 : fz0) + ((c8 & 1) + (int64_t)fn.x * ((c8 >> 1) & 1));
    sup[c8] = s.motion[vb + f] > .5f;
    v[c8] = forward ? s.motion[fb + f] : getv(s, acc, ub + f);
  }
  lim_lo = 1e20f; lim_hi = -1e20f; valid = true;
#pragma unroll
  for (int c8 = 0; c8 < 8; c8++) corner_update(at.t, c8, v[c8], sup[c8], valid, lim_lo, lim_hi);
  F3 t = at.t;
  float a = fmaf(t.x, v[1] - v[0], v[0]), b = fmaf(t.x, v[3] - v[2], v[2]);
  float d = fmaf(t.x, v[5] - v[4], v[4]), e = fmaf(t.x, v[7] - v[6], v[6]);
  float lower = fmaf(t.y, b - a, a), upper = fmaf(t.y, e - d, d);
  return fmaf(t.z, upper - lower, lower);
}
__device__ __forceinline__ float sample_general(const Ctx& s, const SampleAt& at, int axis, bool forward, bool& valid, float& lim_lo, float& lim_hi) {
  I3 lo = at.lo; F3 t = at.t;
  lim_lo = 1e20f; lim_hi = -1e20f; valid = true;
  I3 fn = dims(s); addc(fn, axis, 1);
  int acc = accepted(s);
  float a = 0, b = 0, d = 0, e = 0;
#pragma unroll 1
  for (int pr = 0; pr < 4; pr++) {
    int y = pr & 1, z = pr >> 1;
    bool s0, s1;
    float v0 = sample_tap(s, lo, 0, y, z, axis, forward, at.direct, fn, acc, s0);
    float v1 = sample_tap(s, lo, 1, y, z, axis, forward, at.direct, fn, acc, s1);
    corner_update(t, 2 * pr, v0, s0, valid, lim_lo, lim_hi);
    corner_update(t, 2 * pr + 1, v1, s1, valid, lim_lo, lim_hi);
    float r = fmaf(t.x, v1 - v0, v0);
    a = pr == 0 ? r : a; b = pr == 1 ? r : b; d = pr == 2 ? r : d; e = pr == 3 ? r : e;
  }
  float lower = fmaf(t.y, b - a, a), upper = fmaf(t.y, e - d, d);
  return fmaf(t.z, upper - lower, lower);
}
// The general sample with eight lanes per face (the deferred MacCormack
// faces): lane `sub` of the group evaluates corner `sub` through the same
// donor code, the eight values and supports are gathered by 8-wide shuffles,
// and every lane then runs the corner order and the fma interpolation of the
// single-lane form, so all lanes hold the same, identical result.
__device__ __forceinline__ float sample_general_coop(const Ctx& s, const SampleAt& at, int axis, bool forward, bool& valid, float& lim_lo, float& lim_hi, int sub) {
  I3 fn = dims(s); addc(fn, axis, 1);
  int acc = accepted(s);
  bool mine_s;
  float mine = sample_tap(s, at.lo, sub & 1, (sub >> 1) & 1, sub >> 2, axis, forward, at.direct, fn, acc, mine_s);
  float v[8]; bool sup[8];
#pragma unroll
  for (int c8 = 0; c8 < 8; c8++) { v[c8] = __shfl(mine, c8, 8); sup[c8] = __shfl((int)mine_s, c8, 8) != 0; }
  lim_lo = 1e20f; lim_hi = -1e20f; valid = true;
#pragma unroll
  for (int c8 = 0; c8 < 8; c8++) corner_update(at.t, c8, v[c8], sup[c8], valid, lim_lo, lim_hi);
  F3 t = at.t;
  float a = fmaf(t.x, v[1] - v[0], v[0]), b = fmaf(t.x, v[3] - v[2], v[2]);
  float d = fmaf(t.x, v[5] - v[4], v[4]), e = fmaf(t.x, v[7] - v[6], v[6]);
  float lower = fmaf(t.y, b - a, a), upper = fmaf(t.y, e - d, d);
  return fmaf(t.z, upper - lower, lower);
}
__device__ __forceinline__ float sampled_component(const Ctx& s, F3 x, int axis, bool forward, bool& valid, float& lim_lo, float& lim_hi) {
  SampleAt at = sample_at(s, x, axis);
  return sample_general(s, at, axis, forward, valid, lim_lo, lim_hi);
}
// The samples of one RK2 trace of step h from x (coupled_motion.glslinc:168-188):
// velocity at x (or the cached sv), at mid = x - 0.5 h u(x), and every component
// at end = x - h u(mid), `component`'s from the forward field if forward_end.
// A stage whose three samples are all direct issues their 48 loads together;
// otherwise its samples run one at a time through the general path.  With
// DEFER the trace stops at the first stage that would need the general path
// and reports `deferred` (the caller queues the face for a second pass that
// traces it again in full, so only those faces' waves run the general code).
struct Rk2Trace { F3 sv, um, mid, end; float value, l0, l1; bool vs, vm, ve, deferred; };
template <bool DEFER = false, bool COOP = false>
__device__ __forceinline__ Rk2Trace rk2_samples(const Ctx& s, F3 x, float h, int component, bool forward_end, bool have_sv, F3 sv, int sub = 0) {
  Rk2Trace t;
  t.sv = sv; t.um = mkf(0, 0, 0); t.end = x; t.value = 0; t.l0 = 0; t.l1 = 0; t.vs = true; t.vm = true; t.ve = true; t.deferred = false;
  t.mid = have_sv ? mkf(x.x - .5f * h * sv.x, x.y - .5f * h * sv.y, x.z - .5f * h * sv.z) : x;
#pragma unroll 1
  for (int stage = have_sv ? 1 : 0; stage < 3; stage++) {
    F3 p = stage == 0 ? x : (stage == 1 ? t.mid : t.end);
    bool f0 = stage == 2 && component == 0 && forward_end, f1 = stage == 2 && component == 1 && forward_end,
         f2 = stage == 2 && component == 2 && forward_end;
    SampleAt a0 = sample_at(s, p, 0), a1 = sample_at(s, p, 1), a2 = sample_at(s, p, 2);
    float v0, v1, v2, lo0, lo1, lo2, hi0, hi1, hi2; bool ok0, ok1, ok2;
    if (DEFER && !(a0.direct && a1.direct && a2.direct)) { t.deferred = true; return t; }
    if (a0.direct && a1.direct && a2.direct) {
      v0 = sample_direct(s, a0, 0, f0, ok0, lo0, hi0);
      v1 = sample_direct(s, a1, 1, f1, ok1, lo1, hi1);
      v2 = sample_direct(s, a2, 2, f2, ok2, lo2, hi2);
    } else if (COOP) {
      v0 = sample_general_coop(s, a0, 0, f0, ok0, lo0, hi0, sub);
      v1 = sample_general_coop(s, a1, 1, f1, ok1, lo1, hi1, sub);
      v2 = sample_general_coop(s, a2, 2, f2, ok2, lo2, hi2, sub);
    } else {
      v0 = sample_general(s, a0, 0, f0, ok0, lo0, hi0);
      v1 = sample_general(s, a1, 1, f1, ok1, lo1, hi1);
      v2 = sample_general(s, a2, 2, f2, ok2, lo2, hi2);
    }
    bool ok = ok0 && ok1 && ok2;
    if (stage == 0) { t.sv = mkf(v0, v1, v2); t.vs = ok; t.mid = mkf(x.x - .5f * h * t.sv.x, x.y - .5f * h * t.sv.y, x.z - .5f * h * t.sv.z); }
    else if (stage == 1) { t.um = mkf(v0, v1, v2); t.vm = ok; t.end = mkf(x.x - h * t.um.x, x.y - h * t.um.y, x.z - h * t.um.z); }
    else {
      t.ve = ok;
      t.value = component == 0 ? v0 : (component == 1 ? v1 : v2);
      t.l0 = component == 0 ? lo0 : (component == 1 ? lo1 : lo2);
      t.l1 = component == 0 ? hi0 : (component == 1 ? hi1 : hi2);
    }
  }
  return t;
}
__device__ __forceinline__ F3 sampled_velocity_checked(const Ctx& s, F3 x, bool& supported) {
  F3 v; supported = true;
  for (int axis = 0; axis < 3; axis++) { bool ok; float a, b; setc(v, axis, sampled_component(s, x, axis, false, ok, a, b)); supported = supported && ok; }
  return v;
}
__device__ __forceinline__ float checked_endpoint(const Ctx& s, F3 x, int component, bool forward, bool& supported, float& l0, float& l1) {
  float value = 0; supported = true; l0 = 0; l1 = 0;
  for (int axis = 0; axis < 3; axis++) {
    bool ok; float a, b; float dv = sampled_component(s, x, axis, axis == component && forward, ok, a, b);
    supported = supported && ok; if (axis == component) { value = dv; l0 = a; l1 = b; }
  }
  return value;
}
__device__ __forceinline__ I3 cell_of(const Ctx& s, F3 p) { return mk((int)floorf(p.x / s.h[0]), (int)floorf(p.y / s.h[1]), (int)floorf(p.z / s.h[2])); }
__device__ __forceinline__ bool trace_points_clear(const Ctx& s, F3 x, F3 mid, F3 end) {
  return !solid(s, cell_of(s, x)) && !solid(s, cell_of(s, mid)) && !solid(s, cell_of(s, end));
}
__device__ __forceinline__ F3 trace_rk2(const Ctx& s, F3 x, float dt) {
  bool ok; F3 u = sampled_velocity_checked(s, x, ok);
  F3 mid = mkf(x.x - .5f * dt * u.x, x.y - .5f * dt * u.y, x.z - .5f * dt * u.z);
  u = sampled_velocity_checked(s, mid, ok);
  return mkf(x.x - dt * u.x, x.y - dt * u.y, x.z - dt * u.z);
}
__device__ __forceinline__ bool trace_supported(const Ctx& s, F3 x, float dt) {
  bool ok; F3 u = sampled_velocity_checked(s, x, ok);
  F3 mid = mkf(x.x - .5f * dt * u.x, x.y - .5f * dt * u.y, x.z - .5f * dt * u.z);
  u = sampled_velocity_checked(s, mid, ok);
  F3 end = mkf(x.x - dt * u.x, x.y - dt * u.y, x.z - dt * u.z);
  for (int stage = 0; stage < 3; stage++) {
    F3 p = stage == 0 ? x : (stage == 1 ? mid : end);
    if (solid(s, cell_of(s, p))) return false;
    for (int axis = 0; axis < 3; axis++) { bool v; float a, b; sampled_component(s, p, axis, false, v, a, b); if (!v) return false; }
  }
  return true;
}
// coupled_motion.glslinc:189-202
__device__ __forceinline__ float courant_at(const Ctx& s, I3 q, int bank, bool accelerate) {
  F3 speed;
  for (int axis = 0; axis < 3; axis++) {
    I3 r = q; addc(r, axis, 1);
    setc(speed, axis, gmax(fabsf(getv(s, bank, face_index(s, q, axis))), fabsf(getv(s, bank, face_index(s, r, axis)))));
  }
  if (accelerate && coastal_rate(s, center(s, q).x) > 0) {
    float se, ss; coastal_surface(s, center(s, q).x, s.c[1] + s.dt_max, se, ss);
    if (center(s, q).y < se + 3 * s.h[1]) {
      for (int axis = 0; axis < 3; axis++) {
        I3 r = q; addc(r, axis, 1);
        setc(speed, axis, gmax(compc(speed, axis), gmax(fabsf(coastal_velocity_axis(s, coastal_face_position(s, q, axis), s.c[1] + s.dt_max, axis)),
                                                          fabsf(coastal_velocity_axis(s, coastal_face_position(s, r, axis), s.c[1] + s.dt_max, axis)))));
      }
    }
  }
  if (accelerate) speed.y += 9.81f * s.dt_max * s.gmult;
  speed.x /= s.h[0]; speed.y /= s.h[1]; speed.z /= s.h[2];
  return s.max_face_cfl > .5f ? gmax(speed.x, gmax(speed.y, speed.z)) : dot3(speed, mkf(1, 1, 1));
}
__device__ __forceinline__ float linear_streamfunction(const Ctx& s, F3 x) {
  float A = s.wave[0], k = s.wave[1], w = s.wave[2], h = s.wave[3];
  float value = A * w / (k * sinhf(k * h)) * sinhf(k * (x.y + h)) * cosf(k * x.x);
  if (s.stokes2 > .5f) value += .375f * A * A * w / powf(sinhf(k * h), 4.0f) * sinhf(2 * k * (x.y + h)) * cosf(2 * k * x.x);
  return value;
}
__device__ __forceinline__ bool projected_face(const Ctx& s, I3 q, int axis) {
  q = wrap_face(s, q); if (!face_in_bounds(s, q, axis) || wall_face(s, q, axis)) return false;
  I3 l = q; addc(l, axis, -1); return wet(s, q) || wet(s, l);
}
// coupled_motion.glslinc:212-230 (linear air continuation, numerics.y)
__device__ __forceinline__ float extended_value(const Ctx& s, int64_t source, I3 target) {
  int cb = candidate(s);
  float value = getv(s, cb, U_(s) + source);
  if (s.linear_air < .5f) return value;
  I3 q; int component; unpack_face(s, source, q, component);
  F3 g = mkf(0, 0, 0);
  for (int axis = 0; axis < 3; axis++) {
    I3 l = q, r = q; addc(l, axis, -1); addc(r, axis, 1); l = wrap_face(s, l); r = wrap_face(s, r);
    bool left = projected_face(s, l, component), right = projected_face(s, r, component);
    float dl = left ? value - getv(s, cb, face_index(s, l, component)) : 0;
    float dr = right ? getv(s, cb, face_index(s, r, component)) - value : 0;
    if (left && right) setc(g, axis, dl * dr > 0 ? gsign(dl) * gmin(fabsf(dl), fabsf(dr)) : 0);
    else setc(g, axis, left ? dl : (right ? dr : 0));
  }
  F3 off = periodic_delta(s, target, q);
  off = mkf(gclamp(off.x / s.h[0], -1, 1), gclamp(off.y / s.h[1], -1, 1), gclamp(off.z / s.h[2], -1, 1));
  return value + dot3(g, off);
}
