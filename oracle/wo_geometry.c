/* PLIC, interface geometry and coastal target of the oracle (test infrastructure).
 * Restates water/shaders/compute/plic_math.glslinc, pressure_geometry.glslinc,
 * height_normal.glslinc and coastal_target.glslinc. */
#include "wo_state.h"

/* plic_math.glslinc:3-23 */
float wo_cube_volume3(F3 normal, float alpha) {
  alpha -= fminf_(normal.v[0], 0.0f) + fminf_(normal.v[1], 0.0f) + fminf_(normal.v[2], 0.0f);
  F3 v = f3(fabsf(normal.v[0]), fabsf(normal.v[1]), fabsf(normal.v[2]));
  float total = v.v[0] + v.v[1] + v.v[2];
  if (alpha <= 0.0f) return 0.0f;
  if (alpha >= total) return 1.0f;
  float t;
  if (v.v[0] > v.v[1]) { t = v.v[0]; v.v[0] = v.v[1]; v.v[1] = t; }
  if (v.v[1] > v.v[2]) { t = v.v[1]; v.v[1] = v.v[2]; v.v[2] = t; }
  if (v.v[0] > v.v[1]) { t = v.v[0]; v.v[0] = v.v[1]; v.v[1] = t; }
  int complement = alpha > 0.5f * total;
  t = complement ? total - alpha : alpha;
  float x = v.v[0], y = v.v[1], z = v.v[2], result;
  if (x < 1e-6f * total) {
    t -= x * 0.5f;
    if (y < 1e-6f * total) result = (t - y * 0.5f) / z;
    else if (t < y) result = t * t / (2 * y * z);
    else result = (t - 0.5f * y) / z;
  } else if (t < x) result = t * t * t / (6 * x * y * z);
  else if (t < y) result = (3 * t * t - 3 * t * x + x * x) / (6 * y * z);
  else if (x + y < z && t >= x + y) result = (t - 0.5f * (x + y)) / z;
  else {
    /* GLSL pow(v, 3.0) carries only exp2/log2 precision; the cube is the
       product (v * v) * v here and on the device */
    float ty = fmaxf_(t - y, 0.0f), tz = fmaxf_(t - z, 0.0f);
    float numerator = x * (3 * t * t - 3 * t * x + x * x) - ty * ty * ty - tz * tz * tz;
    result = numerator / (6 * x * y * z);
  }
  return complement ? 1.0f - result : result;
}

/* plic_math.glslinc:24-37 */
float wo_inverse_volume3(F3 normal, float fraction) {
  float lo = fminf_(normal.v[0], 0.0f) + fminf_(normal.v[1], 0.0f) + fminf_(normal.v[2], 0.0f);
  float hi = fmaxf_(normal.v[0], 0.0f) + fmaxf_(normal.v[1], 0.0f) + fmaxf_(normal.v[2], 0.0f);
  for (int it = 0; it < 25; it++) { float mid = 0.5f * (lo + hi); if (wo_cube_volume3(normal, mid) < fraction) lo = mid; else hi = mid; }
  float alpha = 0.5f * (lo + hi), ctr = dot3(normal, f3(.5f, .5f, .5f));
  if (fraction == .5f) return ctr;
  if (fraction < .5f) return fminf_(alpha, ctr);
  if (alpha <= ctr) alpha = ctr + 5.96046448e-8f * dot3(f3(fabsf(normal.v[0]), fabsf(normal.v[1]), fabsf(normal.v[2])), f3(1, 1, 1));
  return alpha;
}

float wo_cube_volume(const float *n, float alpha) { return wo_cube_volume3(f3(n[0], n[1], n[2]), alpha); }
float wo_inverse_volume(const float *n, float f) { return wo_inverse_volume3(f3(n[0], n[1], n[2]), f); }

/* ---- coastal_target.glslinc ---- */
float wo_smooth_ramp(float q) { q = clampf(q, 0, 1); return q * q * q * (10 + q * (-15 + 6 * q)); }
static float ramped_amplitude(const WO *s, float time) {
  return PC_WAVE(s, 0) * wo_smooth_ramp(time * PC_WAVE(s, 2) / (6.28318530718f * PC_FORC(s, 3)));
}
void wo_coastal_surface(const WO *s, float x, float time, float out2[2]) {
  float a = ramped_amplitude(s, time), k = PC_WAVE(s, 1), w = PC_WAVE(s, 2), h = PC_WAVE(s, 3);
  float sigma = tanhf(k * h), second = k * a * a * (3 - sigma * sigma) / (4 * sigma * sigma * sigma);
  if (PC_NUM(s, 2) < .5f) second = 0;
  float phase = k * (x - PC_FORC(s, 0)) - w * time;
  out2[0] = a * cosf(phase) + second * cosf(2 * phase);
  out2[1] = -k * a * sinf(phase) - 2 * k * second * sinf(2 * phase);
}
F3 wo_coastal_velocity(const WO *s, F3 world, float time) {
  float a = ramped_amplitude(s, time), k = PC_WAVE(s, 1), w = PC_WAVE(s, 2), h = PC_WAVE(s, 3);
  float sh = sinhf(k * h), phase = k * (world.v[0] - PC_FORC(s, 0)) - w * time, y = world.v[1] + h;
  float first = a * w / sh, second = PC_NUM(s, 2) > .5f ? .75f * a * a * w * k / powf(sh, 4.0f) : 0;
  return f3(first * coshf(k * y) * cosf(phase) + second * coshf(2 * k * y) * cosf(2 * phase),
            first * sinhf(k * y) * sinf(phase) + second * sinhf(2 * k * y) * sinf(2 * phase), 0);
}
F3 wo_coastal_face_position(const WO *s, I3 q, int axis) { F3 p = center(s, q); p.v[axis] -= .5f * PC_H(s, axis); return p; }
F4 wo_coastal_plane(const WO *s, I3 q, float time) {
  float surf[2]; wo_coastal_surface(s, center(s, q).v[0], time, surf);
  F3 normal = f3(-surf[1] * PC_H(s, 0), PC_H(s, 1), 0);
  return f4(normal.v[0], normal.v[1], normal.v[2], dot3(normal, f3(.5f, .5f, .5f)) + surf[0] - center(s, q).v[1]);
}
float wo_coastal_fraction(const WO *s, I3 q, float time) {
  F4 eq = wo_coastal_plane(s, q, time);
  return clampf(wo_cube_volume3(f3(eq.v[0], eq.v[1], eq.v[2]), eq.v[3]), 0, 1);
}
float wo_coastal_rate(const WO *s, float x) {
  float left = 1 - wo_smooth_ramp((x - PC_FORC(s, 0)) / PC_FORC(s, 1));
  float right = PC_FIX(s, 0) == 6 ? wo_smooth_ramp((x - (PC_O(s, 0) + (float)s->nx * PC_H(s, 0) - PC_FORC(s, 1))) / PC_FORC(s, 1)) : 0;
  return PC_FORC(s, 2) * fmaxf_(left, s->boundary_options[0] * right);
}

/* pressure_geometry.glslinc:39-47 */
float wo_geometry_fraction(const WO *s, I3 q) {
  q = wrap_cell(s, q);
  if (q.v[0] < 0 && !solid(s, q)) return wo_coastal_fraction(s, q, s->c[1] + (geometry_bank(s) == candidate(s) ? timestep(s) : 0));
  if (solid(s, q)) { for (int k = 0; k < 3; k++) q.v[k] = clampi(q.v[k], 0, dims(s).v[k] - 1); }
  if (outside(s, q)) return 0.0f;
  return get_value(s, geometry_bank(s), index3(q, dims(s)));
}
/* coupled_motion.glslinc:36-44 */
float wo_transport_f(const WO *s, I3 q) {
  q = wrap_cell(s, q);
  if (q.v[0] < 0 && !solid(s, q)) return wo_coastal_fraction(s, q, s->c[1]);
  if (outside(s, q) || solid(s, q)) return 0;
  int64_t i = index3(q, dims(s)); int sweep = (int)s->c[21];
  return sweep == 0 ? get_value(s, accepted(s), i) : (sweep == 1 ? get_value(s, candidate(s), i) : s->motion[i]);
}
static float hf_fraction(const WO *s, I3 q, int source) { return source == 1 ? wo_transport_f(s, q) : wo_geometry_fraction(s, q); }
static int hf_valid(const WO *s, I3 q) { q = wrap_cell(s, q); return !outside(s, q) && !solid(s, q); }
static int height_normals_enabled(const WO *s) { return ((int)PC_NUM(s, 3) & 1) != 0; }

/* pressure_geometry.glslinc:72-90 */
F3 wo_interface_gradient(const WO *s, I3 q, int source) {
  F3 g = f3(0, 0, 0);
  for (int axis = 0; axis < 3; axis++) {
    int b = (axis + 1) % 3, d = (axis + 2) % 3, weights = 0;
    for (int j = -1; j <= 1; j++) for (int k = -1; k <= 1; k++) {
      I3 middle = q; middle.v[b] += j; middle.v[d] += k; if (solid(s, middle)) continue;
      I3 l = middle, r = middle; l.v[axis]--; r.v[axis]++;
      int left = !solid(s, l), right = !solid(s, r); if (!left && !right) continue;
      float difference;
      if (left && right) difference = hf_fraction(s, l, source) - hf_fraction(s, r, source);
      else if (left) difference = 2 * (hf_fraction(s, l, source) - hf_fraction(s, middle, source));
      else difference = 2 * (hf_fraction(s, middle, source) - hf_fraction(s, r, source));
      int weight = (j == 0 ? 2 : 1) * (k == 0 ? 2 : 1); weights += weight;
      g.v[axis] += (float)weight * difference / (32 * PC_H(s, axis));
    }
    if (weights > 0) g.v[axis] *= 16.0f / (float)weights;
  }
  return g;
}

/* height_normal.glslinc:6-34 */
F3 wo_height_normal(const WO *s, I3 q, F3 youngs, int source, int *used) {
  *used = 0; F3 scaled;
  for (int k = 0; k < 3; k++) scaled.v[k] = fabsf(youngs.v[k]) * PC_H(s, k);
  int axis = scaled.v[0] > scaled.v[1] ? 0 : 1; if (scaled.v[2] > scaled.v[axis]) axis = 2;
  if (scaled.v[axis] < 1e-12f) return youngs;
  int b = (axis + 1) % 3, d = (axis + 2) % 3; float orientation = signf(youngs.v[axis]);
  float heights[5];
  for (int column = 0; column < 5; column++) {
    I3 base = q;
    if (column == 1) base.v[b]--; if (column == 2) base.v[b]++;
    if (column == 3) base.v[d]--; if (column == 4) base.v[d]++;
    float sum = 0, prev = orientation > 0 ? 1 : 0;
    for (int t = -4; t <= 4; t++) {
      I3 cell = base; cell.v[axis] += t; if (!hf_valid(s, cell)) return youngs;
      float f = hf_fraction(s, cell, source);
      if (isnan(f) || isinf(f) || f < 0 || f > 1) return youngs;
      if (orientation * (f - prev) > 2e-6f) return youngs;
      if (t == -4 && f != (orientation > 0 ? 1 : 0)) return youngs;
      if (t == 4 && f != (orientation > 0 ? 0 : 1)) return youngs;
      sum += f; prev = f;
    }
    heights[column] = sum * PC_H(s, axis);
  }
  F3 r = f3(0, 0, 0); r.v[axis] = orientation;
  r.v[b] = -(heights[2] - heights[1]) / (2 * PC_H(s, b));
  r.v[d] = -(heights[4] - heights[3]) / (2 * PC_H(s, d));
  *used = 1; return r;
}

static F4 plane_from_gradient(const WO *s, I3 q, F3 g, int source, float f) {
  if (height_normals_enabled(s)) { int used; g = wo_height_normal(s, q, g, source, &used); }
  if (len3(g) < 1e-9f) g = f3(0, 1, 0);
  float L = len3(g); F3 normal;
  for (int k = 0; k < 3; k++) normal.v[k] = g.v[k] / L * PC_H(s, k);
  float l1 = dot3(f3(fabsf(normal.v[0]), fabsf(normal.v[1]), fabsf(normal.v[2])), f3(1, 1, 1));
  for (int k = 0; k < 3; k++) normal.v[k] /= l1;
  return f4(normal.v[0], normal.v[1], normal.v[2], wo_inverse_volume3(normal, f));
}
/* pressure_geometry.glslinc:91-98 */
F4 wo_reconstruct_pressure(const WO *s, I3 q) {
  float f = wo_geometry_fraction(s, q); if (f <= 0 || f >= 1) return f4(0, 1, 0, f);
  return plane_from_gradient(s, q, wo_interface_gradient(s, q, 0), 0, f);
}
/* coupled_motion.glslinc:50-56 */
F4 wo_transport_plane(const WO *s, I3 q) {
  float f = wo_transport_f(s, q); if (f <= 0 || f >= 1) return f4(0, 1, 0, f);
  return plane_from_gradient(s, q, wo_interface_gradient(s, q, 1), 1, f);
}

/* pressure_geometry.glslinc:99-101 */
static float segment_distance(F3 p, F3 a, F3 b) {
  F3 ab, pa, d; for (int k = 0; k < 3; k++) { ab.v[k] = b.v[k] - a.v[k]; pa.v[k] = p.v[k] - a.v[k]; }
  float t = clampf(dot3(pa, ab) / fmaxf_(dot3(ab, ab), 1e-30f), 0, 1);
  for (int k = 0; k < 3; k++) d.v[k] = p.v[k] - (a.v[k] + t * ab.v[k]);
  return len3(d);
}
/* pressure_geometry.glslinc:105-120 */
static float polygon_distance(const WO *s, F3 point, F4 eq) {
  F3 h = f3(PC_H(s, 0), PC_H(s, 1), PC_H(s, 2)), normal;
  for (int k = 0; k < 3; k++) normal.v[k] = eq.v[k] / h.v[k];
  float norm2 = dot3(normal, normal), signed_d = (dot3(normal, point) - eq.v[3]) / sqrtf(norm2);
  F3 proj; float step = dot3(normal, point) - eq.v[3];
  for (int k = 0; k < 3; k++) proj.v[k] = point.v[k] - normal.v[k] * step / norm2;
  if (proj.v[0] >= 0 && proj.v[1] >= 0 && proj.v[2] >= 0 && proj.v[0] <= h.v[0] && proj.v[1] <= h.v[1] && proj.v[2] <= h.v[2]) return fabsf(signed_d);
  F3 vert[12]; int count = 0;
  for (int axis = 0; axis < 3; axis++) for (int j = 0; j < 2; j++) for (int k = 0; k < 2; k++) {
    F3 a = f3(0, 0, 0); a.v[(axis + 1) % 3] = (float)j; a.v[(axis + 2) % 3] = (float)k;
    float den = eq.v[axis]; if (fabsf(den) < 1e-15f) continue;
    float t = (eq.v[3] - dot3(f3(eq.v[0], eq.v[1], eq.v[2]), a)) / den;
    if (t >= 0 && t <= 1) { a.v[axis] = t; for (int m = 0; m < 3; m++) a.v[m] *= h.v[m]; vert[count++] = a; }
  }
  float best = 1e10f;
  for (int j = 0; j < count; j++) {
    F3 d; for (int m = 0; m < 3; m++) d.v[m] = point.v[m] - vert[j].v[m];
    best = fminf_(best, len3(d));
    for (int k = 0; k < j; k++) best = fminf_(best, segment_distance(point, vert[j], vert[k]));
  }
  return best;
}
/* pressure_geometry.glslinc:124-132 */
static int clip_polygon(F3 *poly, int count, F4 eq) {
  F3 result[12]; int size = 0; F3 n = f3(eq.v[0], eq.v[1], eq.v[2]);
  for (int j = 0; j < count; j++) {
    F3 a = poly[j], b = poly[(j + 1) % count]; float da = dot3(n, a) - eq.v[3], db = dot3(n, b) - eq.v[3];
    if (da <= 0) result[size++] = a;
    if ((da < 0 && db > 0) || (da > 0 && db < 0)) {
      float t = da / (da - db); F3 m; for (int k = 0; k < 3; k++) m.v[k] = a.v[k] * (1 - t) + b.v[k] * t;
      result[size++] = m;
    }
  }
  for (int j = 0; j < size; j++) poly[j] = result[j];
  return size;
}
/* pressure_geometry.glslinc:133-156 */
static float transition_distance(const WO *s, F3 point, I3 left, int axis) {
  I3 right = left; right.v[axis]++;
  I3 l = wrap_cell(s, left), r = wrap_cell(s, right);
  if (outside(s, l) || outside(s, r) || solid(s, l) || solid(s, r)) return 1e10f;
  float fl = wo_geometry_fraction(s, l), fr = wo_geometry_fraction(s, r);
  if ((fl <= 0 && fr <= 0) || (fl >= 1 && fr >= 1)) return 1e10f;
  const float *pa = s->plane + 4 * index3(l, dims(s)), *pb = s->plane + 4 * index3(r, dims(s));
  F4 a = f4(pa[0], pa[1], pa[2], pa[3]), b = f4(pb[0], pb[1], pb[2], pb[3]);
  if (fl <= 0) a = f4(0, 0, 0, -1); if (fl >= 1) a = f4(0, 0, 0, 1);
  if (fr <= 0) b = f4(0, 0, 0, -1); if (fr >= 1) b = f4(0, 0, 0, 1);
  b.v[3] += b.v[axis];
  float best = 1e10f; F3 h = f3(PC_H(s, 0), PC_H(s, 1), PC_H(s, 2));
  for (int side = 0; side < 2; side++) {
    F3 poly[12]; int j = (axis + 1) % 3, k = (axis + 2) % 3;
    for (int v = 0; v < 4; v++) { poly[v] = f3(0, 0, 0); poly[v].v[axis] = 1; poly[v].v[j] = (v == 1 || v == 2) ? 1 : 0; poly[v].v[k] = v >= 2 ? 1 : 0; }
    F4 first = side == 0 ? a : b, second = side == 0 ? f4(-b.v[0], -b.v[1], -b.v[2], -b.v[3]) : f4(-a.v[0], -a.v[1], -a.v[2], -a.v[3]);
    int count = clip_polygon(poly, 4, first); count = clip_polygon(poly, count, second);
    if (count == 0) continue;
    F3 projected = point; projected.v[axis] = h.v[axis]; F3 local;
    for (int m = 0; m < 3; m++) local.v[m] = projected.v[m] / h.v[m];
    if (local.v[0] >= 0 && local.v[1] >= 0 && local.v[2] >= 0 && local.v[0] <= 1 && local.v[1] <= 1 && local.v[2] <= 1 &&
        dot3(f3(first.v[0], first.v[1], first.v[2]), local) <= first.v[3] && dot3(f3(second.v[0], second.v[1], second.v[2]), local) <= second.v[3])
      best = fminf_(best, fabsf(point.v[axis] - h.v[axis]));
    for (int v = 0; v < count; v++) {
      F3 p0, p1; for (int m = 0; m < 3; m++) { p0.v[m] = poly[v].v[m] * h.v[m]; p1.v[m] = poly[(v + 1) % count].v[m] * h.v[m]; }
      best = fminf_(best, segment_distance(point, p0, p1));
    }
  }
  return best;
}
/* pressure_geometry.glslinc:157-182 */
float wo_geometry_distance(const WO *s, I3 q) {
  float f = wo_geometry_fraction(s, q); F3 h = f3(PC_H(s, 0), PC_H(s, 1), PC_H(s, 2));
  if (f > 0 && f < 1) {
    const float *e = s->plane + 4 * index3(q, dims(s));
    F3 hn = f3(e[0] / h.v[0], e[1] / h.v[1], e[2] / h.v[2]);
    return (dot3(f3(e[0], e[1], e[2]), f3(.5f, .5f, .5f)) - e[3]) / len3(hn);
  }
  int needs_seed = 0;
  for (int axis = 0; axis < 3; axis++) for (int sign = -1; sign <= 1; sign += 2) {
    I3 nb = q; nb.v[axis] += sign; nb = wrap_cell(s, nb); if (solid(s, nb) || outside(s, nb)) continue;
    float other = wo_geometry_fraction(s, nb); int inside = other >= 1;
    if (other > 0 && other < 1) { const float *e = s->plane + 4 * index3(nb, dims(s)); inside = dot3(f3(e[0], e[1], e[2]), f3(.5f, .5f, .5f)) < e[3]; }
    if (inside != (f >= 1)) needs_seed = 1;
  }
  if (!needs_seed) return (f >= 1 ? -1 : 1) * len3(h) * 4;
  float best = len3(h) * 4;
  for (int x = -1; x <= 1; x++) for (int y = -1; y <= 1; y++) for (int z = -1; z <= 1; z++) {
    I3 nbr = i3(q.v[0] + x, q.v[1] + y, q.v[2] + z), wr = wrap_cell(s, nbr);
    if (outside(s, wr) || solid(s, wr)) continue;
    float fraction = wo_geometry_fraction(s, wr); F3 point, cl, d;
    for (int m = 0; m < 3; m++) { point.v[m] = ((float)(q.v[m] - nbr.v[m]) + .5f) * h.v[m]; cl.v[m] = clampf(point.v[m], 0, h.v[m]); d.v[m] = point.v[m] - cl.v[m]; }
    if (len3(d) >= best) continue;
    if (fraction > 0 && fraction < 1) { const float *e = s->plane + 4 * index3(wr, dims(s)); best = fminf_(best, polygon_distance(s, point, f4(e[0], e[1], e[2], e[3]))); }
    for (int axis = 0; axis < 3; axis++) best = fminf_(best, transition_distance(s, point, nbr, axis));
  }
  return (f >= 1 ? -1 : 1) * best;
}
