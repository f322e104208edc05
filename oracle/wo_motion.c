/* Transport, MacCormack, extension, acceptance-side motion ops and coastal
 * forcing of the oracle (test infrastructure). Restates
 * water/shaders/compute/coupled_motion.glslinc and coastal_pass.glslinc. */
#include "wo_state.h"

/* coupled_motion.glslinc:18-35 */
static F3 periodic_delta(const WO *s, I3 a, I3 b) {
  F3 d; for (int k = 0; k < 3; k++) d.v[k] = (float)(a.v[k] - b.v[k]);
  d.v[2] -= roundf(d.v[2] / (float)s->nz) * (float)s->nz;
  if (PC_PERX(s) > .5f) d.v[0] -= roundf(d.v[0] / (float)s->nx) * (float)s->nx;
  for (int k = 0; k < 3; k++) d.v[k] *= PC_H(s, k);
  return d;
}
static int wall_face(const WO *s, I3 q, int axis) { I3 l = q; l.v[axis]--; return solid(s, q) || solid(s, l); }
static int face_in_bounds(const WO *s, I3 q, int axis) {
  I3 n = dims(s); n.v[axis]++;
  return q.v[0] >= 0 && q.v[1] >= 0 && q.v[2] >= 0 && q.v[0] < n.v[0] && q.v[1] < n.v[1] && q.v[2] < n.v[2];
}
static I3 wrap_face(const WO *s, I3 q, int axis) {
  (void)axis; q.v[2] = pidx(q.v[2], s->nz); if (PC_PERX(s) > .5f) q.v[0] = pidx(q.v[0], s->nx); return q;
}
static float old_face(const WO *s, I3 q, int axis) {
  q = wrap_face(s, q, axis);
  if (q.v[0] < 0 && !wall_face(s, q, axis)) return wo_coastal_velocity(s, wo_coastal_face_position(s, q, axis), s->c[1]).v[axis];
  if (!face_in_bounds(s, q, axis) || wall_face(s, q, axis)) return 0;
  return get_value(s, accepted(s), face_index(s, q, axis));
}
static float face_courant(const WO *s, I3 q, int axis) { return old_face(s, q, axis) * timestep(s) / PC_H(s, axis); }

/* coupled_motion.glslinc:66-85 */
static int reflect_voxel_donor(const WO *s, I3 *q, int component, float *parity) {
  float best = 1e20f; I3 selected = *q; *parity = 1;
  for (int axis = 0; axis < 3; axis++) for (int sign = -1; sign <= 1; sign += 2) {
    for (int layer = 1; layer <= 3; layer++) {
      I3 edge = *q; edge.v[axis] += sign * layer; edge = wrap_face(s, edge, component);
      if (!face_in_bounds(s, edge, component)) break;
      I3 left = edge; left.v[component]--;
      int sl = solid(s, left), sr = solid(s, edge);
      int boundary = axis == component ? (sl != sr) : (!sl && !sr);
      if (!boundary) continue;
      I3 mirror = *q; mirror.v[axis] += sign * (2 * layer - (axis == component ? 0 : 1)); mirror = wrap_face(s, mirror, component);
      float distance = ((float)layer - (axis == component ? 0 : .5f)) * PC_H(s, axis);
      if (face_in_bounds(s, mirror, component) && !wall_face(s, mirror, component) && distance < best) {
        selected = mirror; best = distance; *parity = axis == component ? -1 : 1;
      }
      break;
    }
  }
  *q = selected; return best < 1e19f;
}
/* coupled_motion.glslinc:86-111 */
static float velocity_donor(const WO *s, I3 q, int axis, int forward, int *valid) {
  float parity = 1;
  if (q.v[1] < 0) { q.v[1] = -q.v[1] - (axis == 1 ? 0 : 1); if (axis == 1) parity = -parity; }
  if (PC_PERX(s) < .5f) {
    if (q.v[0] >= s->nx + (axis == 0 ? 1 : 0)) { q.v[0] = 2 * s->nx - q.v[0] - (axis == 0 ? 0 : 1); if (axis == 0) parity = -parity; }
  }
  q = wrap_face(s, q, axis);
  if (q.v[0] < 0 && q.v[1] >= 0 && q.v[1] < s->ny && !wall_face(s, q, axis)) {
    *valid = 1; return parity * wo_coastal_velocity(s, wo_coastal_face_position(s, q, axis), s->c[1]).v[axis];
  }
  if (!face_in_bounds(s, q, axis)) { *valid = 0; return 0; }
  I3 l = q; l.v[axis]--; int left = solid(s, l), right = solid(s, q);
  if (left != right) { *valid = 1; return 0; }
  if (left && right) {
    float reflection;
    if (!reflect_voxel_donor(s, &q, axis, &reflection)) { *valid = 0; return 0; }
    parity *= reflection;
  }
  int64_t f = face_local(s, q, axis); *valid = s->motion[valid_base(s, accepted(s)) + f] > .5f;
  return parity * (forward ? s->motion[forward_base(s) + f] : get_value(s, accepted(s), face_index(s, q, axis)));
}
/* coupled_motion.glslinc:112-137 */
static float sampled_component(const WO *s, F3 x, int axis, int forward, int *valid, float limits[2]) {
  F3 lattice; for (int k = 0; k < 3; k++) lattice.v[k] = x.v[k] / PC_H(s, k) - .5f;
  lattice.v[axis] += .5f;
  I3 lo; F3 t; for (int k = 0; k < 3; k++) { lo.v[k] = (int)floorf(lattice.v[k]); t.v[k] = fractf(lattice.v[k]); }
  limits[0] = 1e20f; limits[1] = -1e20f; *valid = 1;
  int interior = ((int)PC_NUM(s, 3) & 8) != 0 && lo.v[0] >= 1 && lo.v[1] >= 1 && lo.v[2] >= 1 &&
                 lo.v[0] + 1 < s->nx && lo.v[1] + 1 < s->ny && lo.v[2] + 1 < s->nz && !solid(s, i3(lo.v[0] + 1, lo.v[1] - 1, lo.v[2]));
  I3 fn = dims(s); fn.v[axis]++;
  float values[8];
  for (int z = 0; z < 2; z++) for (int y = 0; y < 2; y++) for (int xx = 0; xx < 2; xx++) {
    int d[3] = {xx, y, z}; int supported; float value;
    if (interior) {
      int64_t f = face_local(s, lo, axis) + (xx + (int64_t)fn.v[0] * (y + (int64_t)fn.v[1] * z));
      supported = s->motion[valid_base(s, accepted(s)) + f] > .5f;
      value = forward ? s->motion[forward_base(s) + f] : get_value(s, accepted(s), U_(s) + f);
    } else value = velocity_donor(s, i3(lo.v[0] + xx, lo.v[1] + y, lo.v[2] + z), axis, forward, &supported);
    values[xx + 2 * y + 4 * z] = value;
    float w[3]; for (int k = 0; k < 3; k++) w[k] = (1 - t.v[k]) * (1 - (float)d[k]) + t.v[k] * (float)d[k];
    if (w[0] * w[1] * w[2] > 0) { *valid = *valid && supported; limits[0] = fminf_(limits[0], value); limits[1] = fmaxf_(limits[1], value); }
  }
  float a = fmaf(t.v[0], values[1] - values[0], values[0]), b = fmaf(t.v[0], values[3] - values[2], values[2]);
  float d = fmaf(t.v[0], values[5] - values[4], values[4]), e = fmaf(t.v[0], values[7] - values[6], values[6]);
  float lower = fmaf(t.v[1], b - a, a), upper = fmaf(t.v[1], e - d, d);
  return fmaf(t.v[2], upper - lower, lower);
}
static F3 sampled_velocity(const WO *s, F3 x) {
  F3 v; for (int axis = 0; axis < 3; axis++) { int valid; float lim[2]; v.v[axis] = sampled_component(s, x, axis, 0, &valid, lim); }
  return v;
}
static F3 trace(const WO *s, F3 x, float dt) {
  F3 u = sampled_velocity(s, x), mid, end;
  for (int k = 0; k < 3; k++) mid.v[k] = x.v[k] - .5f * dt * u.v[k];
  u = sampled_velocity(s, mid);
  for (int k = 0; k < 3; k++) end.v[k] = x.v[k] - dt * u.v[k];
  return end;
}
static I3 cell_of(const WO *s, F3 p) { I3 q; for (int k = 0; k < 3; k++) q.v[k] = (int)floorf(p.v[k] / PC_H(s, k)); return q; }
static int trace_supported(const WO *s, F3 x, float dt) {
  F3 u = sampled_velocity(s, x), mid, end;
  for (int k = 0; k < 3; k++) mid.v[k] = x.v[k] - .5f * dt * u.v[k];
  u = sampled_velocity(s, mid);
  for (int k = 0; k < 3; k++) end.v[k] = x.v[k] - dt * u.v[k];
  for (int stage = 0; stage < 3; stage++) {
    F3 point = stage == 0 ? x : (stage == 1 ? mid : end);
    if (solid(s, cell_of(s, point))) return 0;
    for (int axis = 0; axis < 3; axis++) { int valid; float lim[2]; sampled_component(s, point, axis, 0, &valid, lim); if (!valid) return 0; }
  }
  return 1;
}
static F3 sampled_velocity_checked(const WO *s, F3 x, int *supported) {
  F3 value; *supported = 1;
  for (int axis = 0; axis < 3; axis++) { int valid; float lim[2]; value.v[axis] = sampled_component(s, x, axis, 0, &valid, lim); *supported = *supported && valid; }
  return value;
}
static float checked_endpoint(const WO *s, F3 x, int component, int forward, int *supported, float limits[2]) {
  float value = 0; *supported = 1; limits[0] = 0; limits[1] = 0;
  for (int axis = 0; axis < 3; axis++) {
    int valid; float dl[2]; float dv = sampled_component(s, x, axis, axis == component && forward, &valid, dl);
    *supported = *supported && valid; if (axis == component) { value = dv; limits[0] = dl[0]; limits[1] = dl[1]; }
  }
  return value;
}
static int trace_points_clear(const WO *s, F3 x, F3 mid, F3 end) {
  return !solid(s, cell_of(s, x)) && !solid(s, cell_of(s, mid)) && !solid(s, cell_of(s, end));
}
/* coupled_motion.glslinc:168-188 */
static float cached_advection(WO *s, int64_t i, F3 x, int axis, float dt, int correct) {
  int64_t base = trace_cache_base(s), nf = faces_(s);
  if (!correct) {
    int vs, vm, ve; float limits[2];
    F3 sv = sampled_velocity_checked(s, x, &vs), mid, end;
    for (int k = 0; k < 3; k++) mid.v[k] = x.v[k] - .5f * dt * sv.v[k];
    F3 um = sampled_velocity_checked(s, mid, &vm);
    for (int k = 0; k < 3; k++) end.v[k] = x.v[k] - dt * um.v[k];
    float value = checked_endpoint(s, end, axis, 0, &ve, limits);
    s->motion[base + i] = limits[0]; s->motion[base + nf + i] = limits[1];
    s->motion[base + 2 * nf + i] = vs && vm && ve && trace_points_clear(s, x, mid, end) ? 1 : 0;
    for (int k = 0; k < 3; k++) s->motion[base + (3 + k) * nf + i] = sv.v[k];
    return value;
  }
  float forward = s->motion[forward_base(s) + i]; if (s->motion[base + 2 * nf + i] < .5f) return forward;
  F3 sv; for (int k = 0; k < 3; k++) sv.v[k] = s->motion[base + (3 + k) * nf + i];
  int vm, ve; float unused[2];
  float rdt = -dt; F3 mid, end;
  for (int k = 0; k < 3; k++) mid.v[k] = x.v[k] - .5f * rdt * sv.v[k];
  F3 um = sampled_velocity_checked(s, mid, &vm);
  for (int k = 0; k < 3; k++) end.v[k] = x.v[k] - rdt * um.v[k];
  float backward = checked_endpoint(s, end, axis, 1, &ve, unused);
  float value = forward + .5f * (get_value(s, accepted(s), U_(s) + i) - backward);
  return vm && ve && trace_points_clear(s, x, mid, end) && value >= s->motion[base + i] && value <= s->motion[base + nf + i] ? value : forward;
}
/* coupled_motion.glslinc:189-202 */
static float courant_at(const WO *s, I3 q, int bank, int accelerate) {
  F3 speed;
  for (int axis = 0; axis < 3; axis++) { I3 r = q; r.v[axis]++; speed.v[axis] = fmaxf_(fabsf(get_value(s, bank, face_index(s, q, axis))), fabsf(get_value(s, bank, face_index(s, r, axis)))); }
  if (accelerate && wo_coastal_rate(s, center(s, q).v[0]) > 0) {
    float surf[2]; wo_coastal_surface(s, center(s, q).v[0], s->c[1] + PC_DTMAX(s), surf);
    if (center(s, q).v[1] < surf[0] + 3 * PC_H(s, 1)) {
      for (int axis = 0; axis < 3; axis++) {
        I3 r = q; r.v[axis]++;
        speed.v[axis] = fmaxf_(speed.v[axis], fmaxf_(fabsf(wo_coastal_velocity(s, wo_coastal_face_position(s, q, axis), s->c[1] + PC_DTMAX(s)).v[axis]),
                                                    fabsf(wo_coastal_velocity(s, wo_coastal_face_position(s, r, axis), s->c[1] + PC_DTMAX(s)).v[axis])));
      }
    }
  }
  if (accelerate) speed.v[1] += 9.81f * PC_DTMAX(s) * PC_FIX(s, 1);
  for (int k = 0; k < 3; k++) speed.v[k] /= PC_H(s, k);
  return PC_FIX(s, 2) > .5f ? fmaxf_(speed.v[0], fmaxf_(speed.v[1], speed.v[2])) : dot3(speed, f3(1, 1, 1));
}
/* coupled_motion.glslinc:203-207 */
static float linear_streamfunction(const WO *s, F3 x) {
  float A = PC_WAVE(s, 0), k = PC_WAVE(s, 1), w = PC_WAVE(s, 2), h = PC_WAVE(s, 3);
  float value = A * w / (k * sinhf(k * h)) * sinhf(k * (x.v[1] + h)) * cosf(k * x.v[0]);
  if (PC_NUM(s, 2) > .5f) value += .375f * A * A * w / powf(sinhf(k * h), 4.0f) * sinhf(2 * k * (x.v[1] + h)) * cosf(2 * k * x.v[0]);
  return value;
}
static int projected_face(const WO *s, I3 q, int axis) {
  q = wrap_face(s, q, axis); if (!face_in_bounds(s, q, axis) || wall_face(s, q, axis)) return 0;
  I3 l = q; l.v[axis]--; return wet(s, q) || wet(s, l);
}
/* coupled_motion.glslinc:212-230 */
static float extended_value(const WO *s, int64_t source, I3 target) {
  int cand = candidate(s);
  float value = get_value(s, cand, U_(s) + source);
  if (PC_NUM(s, 1) < .5f) return value;
  I3 q; int component; unpack_face(s, source, &q, &component);
  F3 g = f3(0, 0, 0);
  for (int axis = 0; axis < 3; axis++) {
    I3 l = q, r = q; l.v[axis]--; r.v[axis]++; l = wrap_face(s, l, component); r = wrap_face(s, r, component);
    int left = projected_face(s, l, component), right = projected_face(s, r, component);
    float dl = left ? value - get_value(s, cand, face_index(s, l, component)) : 0;
    float dr = right ? get_value(s, cand, face_index(s, r, component)) - value : 0;
    if (left && right) g.v[axis] = dl * dr > 0 ? signf(dl) * fminf_(fabsf(dl), fabsf(dr)) : 0;
    else g.v[axis] = left ? dl : (right ? dr : 0);
  }
  F3 off = periodic_delta(s, target, q);
  for (int k = 0; k < 3; k++) off.v[k] = clampf(off.v[k] / PC_H(s, k), -1, 1);
  return value + dot3(g, off);
}

#define GROUP_REDUCE(mode, ...)                                   \
  WO_PARALLEL for (int64_t g = 0; g < groups; g++) {                           \
    F4 lane[128];                                                  \
    for (int lid = 0; lid < 128; lid++) { int64_t i = g * 128 + lid; __VA_ARGS__; lane[lid] = v; } \
    s->partial[g] = tree_reduce(lane, mode);                       \
  }

void wo_motion_op(WO *s, int op, int64_t groups) {
  I3 n = dims(s); int64_t N = s->N, NF = faces_(s), total = groups * 128; float dt = timestep(s);
  float *c = s->c;
  int setup = c[5] > .5f && c[7] < .5f;
  float cell = PC_H(s, 0) * PC_H(s, 1) * PC_H(s, 2);
  int acc = accepted(s), cand = candidate(s);
  if (op == 40) { /* coupled_motion.glslinc:234-270 */
    float shape = PC_FIX(s, 0);
    WO_PARALLEL for (int64_t i = 0; i < total; i++) {
      if (i < N) {
        I3 q = coord(i, n); F3 lo, hi; for (int k = 0; k < 3; k++) { lo.v[k] = (float)q.v[k] * PC_H(s, k); hi.v[k] = lo.v[k] + PC_H(s, k); }
        float f = get_value(s, acc, i);
        if (shape > .5f && shape < 2.5f) {
          F3 lower = shape < 1.5f ? f3(0, 0, 0) : f3(.22f, .28f, .06f);
          F3 upper = shape < 1.5f ? f3(.28f, .3f, 100) : f3(.42f, .48f, .22f);
          f = 1; for (int k = 0; k < 3; k++) f *= clampf((fminf_(hi.v[k], upper.v[k]) - fmaxf_(lo.v[k], lower.v[k])) / PC_H(s, k), 0, 1);
          if (shape > 1.5f) f += clampf((.08f - lo.v[1]) / PC_H(s, 1), 0, 1);
        }
        if (shape == 4) {
          f = 0;
          for (int block = 0; block < 2; block++) {
            F3 lower = f3(block == 0 ? .08f : .36f, .12f, 0), upper = f3(block == 0 ? .24f : .52f, .32f, 100);
            float o = 1; for (int k = 0; k < 3; k++) o *= clampf((fminf_(hi.v[k], upper.v[k]) - fmaxf_(lo.v[k], lower.v[k])) / PC_H(s, k), 0, 1);
            f += o;
          }
        }
        s->a[i] = f; s->b[i] = f;
      }
      if (i < NF) {
        I3 q; int axis; unpack_face(s, i, &q, &axis);
        float velocity = shape == 3 && axis == 0 && !wall_face(s, q, axis) && center(s, q).v[1] < PC_ETA(s) + 3 * PC_H(s, 1) ? .25f : 0;
        if (shape == 4 && axis == 0 && !wall_face(s, q, axis)) velocity = (float)q.v[0] * PC_H(s, 0) < .3f ? 2 : -2;
        if (shape == 5) {
          F3 x; for (int k = 0; k < 3; k++) x.v[k] = ((float)q.v[k] + .5f) * PC_H(s, k) + PC_O(s, k);
          x.v[axis] -= .5f * PC_H(s, axis);
          if (x.v[1] < eta(s, q) + 3 * PC_H(s, 1) && !wall_face(s, q, axis)) {
            if (axis == 0) { F3 p = x, m = x; p.v[1] += .5f * PC_H(s, 1); m.v[1] -= .5f * PC_H(s, 1); velocity = (linear_streamfunction(s, p) - linear_streamfunction(s, m)) / PC_H(s, 1); }
            if (axis == 1) { F3 p = x, m = x; p.v[0] += .5f * PC_H(s, 0); m.v[0] -= .5f * PC_H(s, 0); velocity = -(linear_streamfunction(s, p) - linear_streamfunction(s, m)) / PC_H(s, 0); }
          }
        }
        s->a[U_(s) + i] = velocity; s->b[U_(s) + i] = velocity;
      }
    }
    c[20] = PC_DTMAX(s); c[31] = PC_FIX(s, 1);
    return;
  }
  if (op == 62) {
    if (c[4] < .5f || c[6] < .5f || c[7] > .5f) return;
    GROUP_REDUCE(RED_SUM, float r = 0; if (i < N) { I3 q = coord(i, n); if (wet(s, q)) r = 1000 / dt * divergence(s, q); } F4 v = f4(r * r, 0, 0, 0));
    return;
  }
  if (op == 66) {
    if (c[4] < .5f || c[6] < .5f || c[7] > .5f) return;
    float r2 = global_reduce(s, 0).v[0];
    c[60] = sqrtf(r2 / fmaxf_(c[59], 1e-30f)); c[62] = pressure_scaled_norm(s, c[59]); c[63] = pressure_scaled_norm(s, r2);
    c[61] = pressure_norm_converged(s, r2, c[59]) ? 1 : 0;
    return;
  }
  if (op == 63) { if (c[57] == 1 && c[7] < .5f) { c[57] = 2; c[6] = 0; c[8] = 0; } return; }
  if (op == 64) {
    if (c[57] == 2 && c[7] < .5f) for (int64_t i = 0; i < N && i < total; i++) { s->work[13 * N + i] = get_value(s, cand, P_(s) + i); put_value(s, cand, P_(s) + i, 0); }
    return;
  }
  if (op == 65) {
    if (c[4] > .5f && c[6] > .5f && c[7] < .5f) for (int64_t i = 0; i < N && i < total; i++) put_value(s, cand, P_(s) + i, get_value(s, cand, P_(s) + i) + s->work[13 * N + i]);
    return;
  }
  if (op == 55) {
    if (!setup) return;
    int axis = ((int)c[2] + (int)c[21]) % 3;
    GROUP_REDUCE(RED_SUM, F4 v = f4(0, 0, 0, 0); if (i < N) {
      I3 q = coord(i, n);
      if (axis == 1 && q.v[1] == n.v[1] - 1) { q.v[1]++; v.v[0] = s->motion[flux_base(s) + face_local(s, q, 1)] * PC_H(s, 0) * PC_H(s, 1) * PC_H(s, 2); }
      if (axis == 0 && q.v[0] == 0) { float flux = s->motion[flux_base(s) + face_local(s, q, 0)] * PC_H(s, 0) * PC_H(s, 1) * PC_H(s, 2); v.v[1] = fmaxf_(flux, 0); v.v[2] = fmaxf_(-flux, 0); }
    });
    return;
  }
  if (op == 56) { if (!setup) return; F4 v = global_reduce(s, 0); c[39] += v.v[0]; c[65] += v.v[1]; c[66] += v.v[2]; return; }
  if (op == 47) {
    if (c[4] < .5f || c[7] > .5f) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(NF, total); i++) {
      I3 q; int axis; unpack_face(s, i, &q, &axis); int64_t canonical = face_index(s, q, axis);
      if (canonical != U_(s) + i) put_value(s, cand, U_(s) + i, get_value(s, cand, canonical));
    }
    return;
  }
  if (op == 48 || op == 49) {
    int initial = op == 49; if (!initial && (c[4] < .5f || c[6] < .5f || c[7] > .5f)) return;
    WO_PARALLEL for (int64_t i = 0; i < total; i++) {
      if (i < N) {
        float flag = wet(s, coord(i, n)) ? 1 : 0;
        if (initial) { s->motion[pressure_mask_base(s, 0) + i] = flag; s->motion[pressure_mask_base(s, 1) + i] = flag; }
        else s->motion[pressure_mask_base(s, cand) + i] = flag;
      }
      if (i < NF) {
        I3 q; int axis; unpack_face(s, i, &q, &axis); I3 l = q; l.v[axis]--;
        float valid = initial ? (!wall_face(s, q, axis) && (wet(s, q) || wet(s, l) || center(s, q).v[1] < eta(s, q) + 3 * PC_H(s, 1)) ? 1 : 0)
                              : (s->motion[source_base(s, 1) + i] >= 0 ? 1 : 0);
        if (initial) { s->motion[valid_base(s, 0) + i] = valid; s->motion[valid_base(s, 1) + i] = valid; }
        else s->motion[valid_base(s, cand) + i] = valid;
      }
    }
    return;
  }
  if (op == 44) {
    if (c[32] > 0 && c[32] <= PC_FIX(s, 3)) { int64_t row = ((int64_t)c[32] - 1) * 4 + 3; float *h = s->history + 4 * row; h[0] = c[1]; h[1] = c[20]; h[2] = c[7] > 0 ? -c[7] : (c[4] < .5f ? 1 : 0); h[3] = c[2]; }
    return;
  }
  if (op == 45) {
    if (!setup) return;
    GROUP_REDUCE(RED_SUM, F4 v = f4(0, 0, 0, 0); if (i < N) {
      I3 q = coord(i, n); float f = get_value(s, acc, i), dv = 0, raw = 0;
      for (int axis = 0; axis < 3; axis++) {
        I3 r = q; r.v[axis]++;
        dv += (old_face(s, r, axis) - old_face(s, q, axis)) / PC_H(s, axis);
        raw += (get_value(s, acc, face_index(s, r, axis)) - get_value(s, acc, face_index(s, q, axis))) / PC_H(s, axis);
      }
      v = f4(f > .5f ? dt * dv * cell : 0, f > .5f && s->motion[pressure_mask_base(s, acc) + i] < .5f ? 1 : 0, f * cell, f > .5f ? dt * raw * cell : 0);
    });
    return;
  }
  if (op == 46) { if (!setup) return; F4 v = global_reduce(s, 0); c[35] = v.v[0]; c[36] = v.v[1]; c[37] = v.v[2]; c[38] = v.v[3]; if (v.v[1] > 0) { c[7] = 22; c[3]++; } return; }
  if (op == 43) {
    int retry = c[7] == 20 || (c[7] >= 10 && c[7] <= 12);
    if (retry && c[29] < 2) { c[33] = c[7]; c[34] = c[20]; c[30] = c[20] * .5f; c[29]++; c[7] = 0; c[4] = 0; c[5] = 0; }
    return;
  }
  if (op == 41) { GROUP_REDUCE(RED_SUM, F4 v = i < N ? f4(get_value(s, acc, i) * PC_H(s, 0) * PC_H(s, 1) * PC_H(s, 2), 0, 0, 0) : f4(0, 0, 0, 0)); return; }
  if (op == 42) { F4 v = global_reduce(s, 0); c[26] = v.v[0]; return; }
  if (op == 20) {
    if (!setup) return;
    GROUP_REDUCE(RED_BOUNDS, float rate = i < N ? courant_at(s, coord(i, n), acc, 1) : 0; F4 v = f4(rate, rate, 0, 0));
    return;
  }
  if (op == 21) {
    if (!setup) return;
    float rate = 0; int64_t G = (N + 127) / 128;
    for (int64_t k = 0; k < G; k++) rate = fmaxf_(rate, s->partial[k].v[1]);
    c[20] = fminf_(PC_DTMAX(s), .45f / fmaxf_(rate, 1e-20f)); if (c[29] > 0) c[20] = fminf_(c[20], c[30]);
    c[21] = 0; c[22] = 0; c[23] = 0; c[39] = 0;
    for (int field = 64; field < 70; field++) c[field] = 0;
    return;
  }
  if (op == 22) {
    if (!setup) return;
    WO_PARALLEL for (int64_t i = 0; i < total; i++) {
      if (i < s->stride) put_value(s, cand, i, get_value(s, acc, i));
      if (i < N) s->motion[N + i] = get_value(s, acc, i) > .5f ? 1 : 0;
    }
    return;
  }
  if (op == 23) {
    if (!setup) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(N, total); i++) { F4 e = wo_transport_plane(s, coord(i, n)); memcpy(s->plane + 4 * i, e.v, 16); }
    return;
  }
  if (op == 24) {
    if (!setup) return;
    int sweep_axis = ((int)c[2] + (int)c[21]) % 3;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(NF, total); i++) {
      I3 q; int axis; unpack_face(s, i, &q, &axis);
      float amount = 0;
      if (axis == sweep_axis && !wall_face(s, q, axis)) {
        float velocity = old_face(s, q, axis), courant = face_courant(s, q, axis); I3 donor = q; if (velocity > 0) donor.v[axis]--;
        donor = wrap_cell(s, donor); float f = wo_transport_f(s, donor), width = fabsf(courant); amount = f * width;
        if (f > 0 && f < 1 && width > 0) {
          F4 eq;
          if (donor.v[0] < 0) eq = wo_coastal_plane(s, donor, c[1]);
          else { const float *e = s->plane + 4 * index3(donor, n); eq = f4(e[0], e[1], e[2], e[3]); }
          F3 low = f3(0, 0, 0), extent = f3(1, 1, 1); extent.v[axis] = width; if (velocity > 0) low.v[axis] = 1 - width;
          F3 ne; for (int k = 0; k < 3; k++) ne.v[k] = eq.v[k] * extent.v[k];
          amount = width * wo_cube_volume3(ne, eq.v[3] - dot3(f3(eq.v[0], eq.v[1], eq.v[2]), low));
        }
        amount *= signf(velocity);
      }
      s->motion[flux_base(s) + i] = amount;
    }
    return;
  }
  if (op == 25) {
    if (!setup) return;
    int axis = ((int)c[2] + (int)c[21]) % 3; int sweep = (int)c[21];
    GROUP_REDUCE(RED_BOUNDS, F4 v = f4(1e20f, -1e20f, 0, 0); if (i < N) {
      I3 q = coord(i, n), r = q; r.v[axis]++;
      float net = s->motion[flux_base(s) + face_local(s, q, axis)] - s->motion[flux_base(s) + face_local(s, r, axis)];
      float correction = s->motion[N + i] * (face_courant(s, r, axis) - face_courant(s, q, axis));
      float increment = net + correction;
      float f = wo_transport_f(s, q) + increment; float change = 0;
      if (f >= -1e-6f && f <= 1 + 1e-6f) change = clampf(f, 0, 1) - f;
      if (sweep == 1) s->motion[i] = f + change; else put_value(s, cand, i, f + change);
      v = f4(f, f, change, fabsf(change)); if (isnan(f) || isinf(f)) { v.v[0] = -1e30f; v.v[1] = 1e30f; }
    });
    return;
  }
  if (op == 26) {
    if (!setup) return;
    F4 v = f4(1e20f, -1e20f, 0, 0); int64_t G = (N + 127) / 128;
    for (int64_t k = 0; k < G; k++) { v.v[0] = fminf_(v.v[0], s->partial[k].v[0]); v.v[1] = fmaxf_(v.v[1], s->partial[k].v[1]); v.v[2] += s->partial[k].v[2]; v.v[3] += s->partial[k].v[3]; }
    uint32_t step = (uint32_t)(int64_t)c[32] - 1u, sweep = (uint32_t)c[21];
    if (step < (uint32_t)PC_FIX(s, 3)) memcpy(s->history + 4 * ((int64_t)step * 4 + sweep), v.v, 16);
    c[22] += v.v[2] * cell; c[23] += v.v[3] * cell;
    if (v.v[0] < -1e-6f || v.v[1] > 1 + 1e-6f) { c[7] = 10 + c[21]; c[3]++; }
    c[21]++;
    return;
  }
  if (op == 27 || op == 28) {
    if (!setup) return;
    int flags = (int)PC_NUM(s, 3);
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(NF, total); i++) {
      I3 q; int axis; unpack_face(s, i, &q, &axis);
      F3 x; for (int k = 0; k < 3; k++) x.v[k] = ((float)q.v[k] + .5f) * PC_H(s, k);
      x.v[axis] -= .5f * PC_H(s, axis);
      float value = 0;
      if (!wall_face(s, q, axis) && ((flags & 2) == 0 || s->motion[valid_base(s, acc) + i] > .5f)) {
        if (flags & 4) value = cached_advection(s, i, x, axis, dt, op == 28);
        else if (op == 27) { int valid; float lim[2]; value = sampled_component(s, trace(s, x, dt), axis, 0, &valid, lim); }
        else {
          int valid, vf; float lim[2], dl[2];
          float backward = sampled_component(s, trace(s, x, -dt), axis, 1, &valid, lim);
          sampled_component(s, trace(s, x, dt), axis, 0, &vf, dl);
          float forward = s->motion[forward_base(s) + i], corrected = forward + .5f * (get_value(s, acc, U_(s) + i) - backward);
          value = valid && vf && trace_supported(s, x, dt) && trace_supported(s, x, -dt) && corrected >= dl[0] && corrected <= dl[1] ? corrected : forward;
        }
      }
      if (op == 27) s->motion[forward_base(s) + i] = value; else put_value(s, cand, U_(s) + i, value);
    }
    return;
  }
  if (op == 30) {
    if (c[4] < .5f || c[6] < .5f || c[7] > .5f) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(NF, total); i++) {
      I3 q; int axis; unpack_face(s, i, &q, &axis); I3 l = q; l.v[axis]--;
      s->motion[source_base(s, 0) + i] = !wall_face(s, q, axis) && (wet(s, q) || wet(s, l)) ? (float)i : -1;
    }
    return;
  }
  if (op >= 31 && op <= 33) {
    if (c[4] < .5f || c[6] < .5f || c[7] > .5f) return;
    int rb = (op - 31) % 2, wb = 1 - rb;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(NF, total); i++) {
      I3 q; int axis; unpack_face(s, i, &q, &axis); float source = s->motion[source_base(s, rb) + i], best = 1e20f;
      if (!wall_face(s, q, axis)) {
        if (source >= 0) { I3 sq; int sa; unpack_face(s, (int64_t)source, &sq, &sa); F3 d = periodic_delta(s, q, sq); best = dot3(d, d); }
        for (int ax = 0; ax < 3; ax++) for (int sign = -1; sign <= 1; sign += 2) {
          I3 nb = q; nb.v[ax] += sign; nb = wrap_face(s, nb, axis);
          if (!face_in_bounds(s, nb, axis) || wall_face(s, nb, axis)) continue;
          float cs = s->motion[source_base(s, rb) + face_local(s, nb, axis)]; if (cs < 0) continue;
          I3 sq; int sa; unpack_face(s, (int64_t)cs, &sq, &sa); F3 d = periodic_delta(s, q, sq); float distance = dot3(d, d);
          if (distance < best || (distance == best && cs < source)) { best = distance; source = cs; }
        }
      } else source = -1;
      s->motion[source_base(s, wb) + i] = source;
    }
    return;
  }
  if (op == 34) {
    if (c[4] < .5f || c[6] < .5f || c[7] > .5f) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(NF, total); i++) {
      I3 q; int axis; unpack_face(s, i, &q, &axis); I3 l = q; l.v[axis]--;
      if (wet(s, q) || wet(s, l)) continue;
      float source = s->motion[source_base(s, 1) + i];
      float value = source < 0 ? 0 : extended_value(s, (int64_t)source, q);
      put_value(s, cand, U_(s) + i, value);
    }
    return;
  }
  if (op == 35) {
    if (c[4] < .5f || c[6] < .5f || c[7] > .5f) return;
    GROUP_REDUCE(RED_BOUNDS, float rate = i < N ? courant_at(s, coord(i, n), cand, 0) * dt : 0; F4 v = f4(rate, rate, 0, 0));
    return;
  }
  if (op == 36) {
    if (c[4] < .5f || c[6] < .5f || c[7] > .5f) return;
    float rate = 0; int64_t G = (N + 127) / 128;
    for (int64_t k = 0; k < G; k++) rate = fmaxf_(rate, s->partial[k].v[1]);
    c[27] = rate;
    return;
  }
}

static F3 cell_velocity(const WO *s, int bank, I3 q) {
  F3 v; for (int axis = 0; axis < 3; axis++) { I3 r = q; r.v[axis]++; v.v[axis] = .5f * (get_value(s, bank, face_index(s, q, axis)) + get_value(s, bank, face_index(s, r, axis))); }
  return v;
}
static int offshore_zone(const WO *s, float x) { return x < PC_FORC(s, 0) + PC_FORC(s, 1); }

/* coastal_pass.glslinc:42-114 */
void wo_coastal_op(WO *s, int op, int64_t groups) {
  I3 n = dims(s); int64_t N = s->N, NF = faces_(s), total = groups * 128; float *c = s->c;
  int setup = c[5] > .5f && c[7] < .5f; float time = c[1] + timestep(s);
  float cell = PC_H(s, 0) * PC_H(s, 1) * PC_H(s, 2);
  int cand = candidate(s);
  if (op == 80) {
    if (!setup) return;
    GROUP_REDUCE(RED_SUM, F4 v = f4(0, 0, 0, 0); if (i < N) {
      I3 q = coord(i, n);
      if (!solid(s, q)) {
        float old = get_value(s, cand, i), rate = wo_coastal_rate(s, center(s, q).v[0]);
        float target = offshore_zone(s, center(s, q).v[0]) ? wo_coastal_fraction(s, q, time) : clampf(.5f - center(s, q).v[1] / PC_H(s, 1), 0, 1);
        float value = old + (1 - expf(-rate * timestep(s))) * (target - old); put_value(s, cand, i, value);
        F3 cv = cell_velocity(s, cand, q);
        v = f4((value - old) * cell, 1000 * (value - old) * cell * cv.v[0], 1000 * (value - old) * cell * cv.v[1], 1000 * (value - old) * cell * cv.v[2]);
      }
    });
    return;
  }
  if (op == 81) { if (!setup) return; F4 ch = global_reduce(s, 0); c[64] += ch.v[0]; for (int axis = 0; axis < 3; axis++) c[67 + axis] += ch.v[axis + 1]; return; }
  if (op == 82) {
    if (!setup) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(NF, total); i++) {
      I3 q; int axis; unpack_face(s, i, &q, &axis);
      F3 position = wo_coastal_face_position(s, q, axis); float old = get_value(s, cand, U_(s) + i), value = old;
      if (!wall_face(s, q, axis)) {
        float target = offshore_zone(s, position.v[0]) ? wo_coastal_velocity(s, position, time).v[axis] : 0;
        float t = 1 - expf(-wo_coastal_rate(s, position.v[0]) * timestep(s));
        value = old * (1 - t) + target * t;
        if (axis == 0 && q.v[0] == 0) value = target;
      }
      put_value(s, cand, U_(s) + i, value); s->motion[forward_base(s) + i] = value - old;
    }
    return;
  }
  if (op == 83) {
    if (!setup) return;
    GROUP_REDUCE(RED_SUM, F4 v = f4(0, 0, 0, 0); if (i < N) {
      I3 q = coord(i, n); F3 ch;
      for (int axis = 0; axis < 3; axis++) { I3 r = q; r.v[axis]++; ch.v[axis] = .5f * (s->motion[forward_base(s) + face_local(s, q, axis)] + s->motion[forward_base(s) + face_local(s, r, axis)]); }
      float m = get_value(s, cand, i) * 1000 * cell;
      v = f4(ch.v[0] * m, ch.v[1] * m, ch.v[2] * m, 0);
    });
    return;
  }
  if (op == 84) { if (!setup) return; F4 ch = global_reduce(s, 0); for (int axis = 0; axis < 3; axis++) c[67 + axis] += ch.v[axis]; return; }
  if (op == 86) {
    for (int i = 0; i < 8; i++) {
      float probe_time = PC_FORC(s, 3) * 6.28318530718f / PC_WAVE(s, 2) * (i < 4 ? .37f : 1);
      F3 point = f3(PC_O(s, 0) + .17f + (float)i * .23f, -PC_WAVE(s, 3) + (float)i * .07f, 0);
      float surf[2]; wo_coastal_surface(s, point.v[0], probe_time, surf); F3 vel = wo_coastal_velocity(s, point, probe_time);
      s->probe[4 * i] = surf[0]; s->probe[4 * i + 1] = surf[1]; s->probe[4 * i + 2] = vel.v[0]; s->probe[4 * i + 3] = vel.v[1];
    }
    return;
  }
  if (op == 87) {
    if (!setup) return;
    GROUP_REDUCE(RED_BOUNDS, float f = i < N ? get_value(s, cand, i) : 0; F4 v = (isnan(f) || isinf(f)) ? f4(-1e30f, 1e30f, 0, 0) : f4(f, f, 0, 0));
    return;
  }
  if (op == 88) {
    if (!setup) return;
    float mn = 1, mx = 0; int64_t G = (N + 127) / 128;
    for (int64_t k = 0; k < G; k++) { mn = fminf_(mn, s->partial[k].v[0]); mx = fmaxf_(mx, s->partial[k].v[1]); }
    c[88] = mn; c[89] = mx; if (mn < 0 || mx > 1) { c[7] = 23; c[3]++; }
    return;
  }
  if (op == 89) {
    if (c[2] <= c[91] || c[2] > PC_FIX(s, 3)) return;
    int64_t row = ((int64_t)c[2] - 1) * 5; float *h = s->chist + 4 * row;
    float vals[20] = {c[1], c[90], c[17], c[72], c[73], c[74], c[40], c[24], c[80], c[81], c[82], c[83], c[84], c[85], c[86], c[87], c[75], c[76], c[77], c[28]};
    memcpy(h, vals, sizeof vals); c[91] = c[2];
    return;
  }
  if (op == 85) {
    for (int i = 0; i < 8; i++) {
      float begin = PC_FORC(s, 0) + PC_FORC(s, 1) + .4f;
      float end = PC_O(s, 0) + (float)n.v[0] * PC_H(s, 0) - (PC_FIX(s, 0) == 6 ? PC_FORC(s, 1) + .4f : 1.0f);
      float t = (float)i / 7; float m = begin * (1 - t) + end * t;
      int x = clampi((int)((m - PC_O(s, 0)) / PC_H(s, 0)), 0, n.v[0] - 1);
      float height = 0;
      for (int z = 0; z < n.v[2]; z++) {
        float column = PC_O(s, 1);
        for (int y = 0; y < n.v[1]; y++) { I3 q = i3(x, y, z); column += (solid(s, q) ? 1 : get_value(s, accepted(s), index3(q, n))) * PC_H(s, 1); }
        height += column;
      }
      c[80 + i] = height / (float)n.v[2];
    }
    return;
  }
}
