/* Pressure ops of the oracle (test infrastructure): restates
 * water/shaders/compute/pressure_pass.glslinc and pressure_iterations.glslinc. */
#include "wo_state.h"

/* pressure_core.glslinc:135 sum_partial / 139-143 pressure_partial */
#define FOR_GROUPS(s, groups) WO_PARALLEL for (int64_t g = 0; g < (groups); g++)

static void pcg_iteration_op(WO *s, int op, int64_t groups) {
  I3 n = dims(s); float dt = timestep(s); int64_t N = s->N;
  if (op == 3) { /* pressure_iterations.glslinc:4-13 */
    if (!pressure_setup(s)) return;
    FOR_GROUPS(s, groups) {
      F4 lane[128];
      for (int lid = 0; lid < 128; lid++) {
        int64_t i = g * 128 + lid; F4 sums = f4(0, 0, 0, 0);
        if (i < N) {
          I3 q = coord(i, n); float r = 0, z = 0, bv = 0, diag = 1;
          if (wet(s, q)) { diag = diagonal(s, q); bv = -1000.0f / dt * divergence(s, q); r = bv - operator_at(s, q, 0); z = r / diag; }
          s->work[i] = r; s->work[N + i] = z; s->work[2 * N + i] = z; s->work[4 * N + i] = diag; s->work[5 * N + i] = bv;
          sums = f4(r * z, r * r, bv * bv, fabsf(r));
        }
        lane[lid] = sums;
      }
      s->partial[g] = tree_reduce(lane, RED_PRESSURE);
    }
    return;
  }
  if (op == 4) { /* pressure_iterations.glslinc:14-19 */
    if (!pressure_setup(s)) return;
    F4 v = global_reduce(s, 1);
    s->c[9] = v.v[0]; s->c[10] = v.v[2]; s->c[13] = sqrtf(v.v[1] / fmaxf_(v.v[2], 1e-30f));
    if (residual_converged(s, v, v.v[2])) s->c[6] = 1;
    if (s->c[58] == 0) s->c[59] = v.v[2];
    s->c[57] = 0;
    return;
  }
  if (op == 5) {
    if (!solving(s)) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(N, groups * 128); i++) { I3 q = coord(i, n); s->work[3 * N + i] = wet(s, q) ? operator_at(s, q, 1) : 0; }
    return;
  }
  if (op == 6) {
    if (!solving(s)) return;
    FOR_GROUPS(s, groups) {
      F4 lane[128];
      for (int lid = 0; lid < 128; lid++) { int64_t i = g * 128 + lid; lane[lid] = i < N ? f4(s->work[2 * N + i] * s->work[3 * N + i], 0, 0, 0) : f4(0, 0, 0, 0); }
      s->partial[g] = tree_reduce(lane, RED_SUM);
    }
    return;
  }
  if (op == 7) {
    if (!solving(s)) return;
    float dotp = global_reduce(s, 0).v[0];
    if (dotp <= 0 || isnan(dotp) || isinf(dotp)) { s->c[7] = 1; return; }
    s->c[11] = s->c[9] / dotp; return;
  }
  if (op == 8) {
    if (!solving(s)) return;
    int cand = candidate(s); int64_t P = P_(s);
    FOR_GROUPS(s, groups) {
      F4 lane[128];
      for (int lid = 0; lid < 128; lid++) {
        int64_t i = g * 128 + lid; F4 sums = f4(0, 0, 0, 0);
        if (i < N) {
          float r = s->work[i] - s->c[11] * s->work[3 * N + i]; float z = r / s->work[4 * N + i];
          put_value(s, cand, P + i, get_value(s, cand, P + i) + s->c[11] * s->work[2 * N + i]);
          s->work[i] = r; s->work[N + i] = z; sums = f4(r * z, r * r, 0, fabsf(r));
        }
        lane[lid] = sums;
      }
      s->partial[g] = tree_reduce(lane, RED_PRESSURE);
    }
    return;
  }
  if (op == 9) {
    if (!solving(s)) return;
    F4 v = global_reduce(s, 1);
    s->c[12] = v.v[0] / fmaxf_(s->c[9], 1e-30f); s->c[9] = v.v[0]; s->c[8]++; s->c[18]++;
    s->c[13] = sqrtf(v.v[1] / fmaxf_(s->c[10], 1e-30f));
    if (isnan(s->c[13]) || isinf(s->c[13])) s->c[7] = 2;
    else if (residual_converged(s, v, s->c[10])) s->c[6] = 1;
    else if (s->c[8] > 4096) s->c[7] = 3;
    return;
  }
  if (op == 10) {
    if (!solving(s)) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(N, groups * 128); i++) s->work[2 * N + i] = s->work[N + i] + s->c[12] * s->work[2 * N + i];
    return;
  }
}

/* The alongshore surface perturbation of the benchmark configurations (not a
 * reference op; the reference initial state is z-invariant,
 * pressure_pass.glslinc:15-17). delta(z) = amplitude * u(z), u in [-1, 1)
 * from a 32-bit integer hash of (seed, z). */
uint32_t wo_hash32(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16; return x;
}
float wo_perturb_unit(uint32_t seed, int z) {
  uint32_t h = wo_hash32(seed * 0x9e3779b9U + (uint32_t)z * 0x85ebca6bU + 0x632be5abU);
  return (float)(h >> 8) * (1.0f / 8388608.0f) - 1.0f;
}

void wo_pressure_op(WO *s, int op, int64_t groups) {
  I3 n = dims(s); int64_t N = s->N; float dt = timestep(s);
  int64_t total = groups * 128;
  if (op == 0) { /* pressure_pass.glslinc:3-19 */
    WO_PARALLEL for (int64_t i = 0; i < total; i++) {
      if (i < s->stride) { s->a[i] = 0; s->b[i] = 0; }
      if (i < 96) s->c[i] = 0;
      if (i < N) {
        I3 q = coord(i, n); float ph = center(s, q).v[1] - eta(s, q);
        int wall = solid(s, q); s->geo[2 * i] = ph; s->geo[2 * i + 1] = wall ? -1.0f : (ph < 0.0f ? 1.0f : 0.0f);
        float f = wall ? 0.0f : clampf(0.5f - ph / PC_H(s, 1), 0.0f, 1.0f); s->a[i] = f; s->b[i] = f;
      }
    }
    return;
  }
  if (op == 90) { /* seeded alongshore perturbation of the still surface */
    float amp = s->opts[6]; uint32_t seed = (uint32_t)s->opts[7];
    if (amp == 0) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(N, total); i++) {
      I3 q = coord(i, n); if (solid(s, q)) continue;
      float ph = center(s, q).v[1] - (eta(s, q) + amp * wo_perturb_unit(seed, q.v[2]));
      float f = clampf(0.5f - ph / PC_H(s, 1), 0.0f, 1.0f); s->a[i] = f; s->b[i] = f;
    }
    return;
  }
  if (op == 1) { /* pressure_pass.glslinc:20-24 */
    float quota = PC_QUOTA(s);
    s->c[5] = 0;
    if (s->c[7] > 0.5f || (quota >= 0 ? s->c[2] >= quota : s->c[1] >= -quota)) return;
    if (s->c[4] < 0.5f) { s->c[4] = 1; s->c[5] = 1; s->c[6] = 0; s->c[8] = 0; s->c[13] = 1; s->c[32]++; s->c[57] = 0; s->c[58] = 0; }
    return;
  }
  if (op == 2) { /* pressure_pass.glslinc:25-50 */
    if (s->c[5] < 0.5f) return;
    int acc = accepted(s), cand = candidate(s); int64_t P = P_(s), Ub = U_(s), Vb = V_(s), Wb = W_(s);
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(total, s->stride); i++) {
      float val = get_value(s, cand, i);
      if (i < N) s->work[13 * N + i] = 0;
      if (i >= P) val = wet(s, coord(i - P, n)) ? get_value(s, acc, i) : 0;
      if (i >= Ub && i < P) {
        int axis = i < Vb ? 0 : (i < Wb ? 1 : 2); int64_t base = axis == 0 ? Ub : (axis == 1 ? Vb : Wb); I3 fn = n; fn.v[axis]++;
        I3 r = coord(i - base, fn), l = r; l.v[axis]--;
        if (solid(s, l) || solid(s, r) || (!wet(s, l) && !wet(s, r))) val = 0.0f;
        else if (axis == 1) val -= 9.81f * dt * s->c[31];
      }
      put_value(s, cand, i, val);
    }
    return;
  }
  if (op >= 3 && op <= 10) { pcg_iteration_op(s, op, groups); return; }
  if (op == 11) { /* pressure_pass.glslinc:52-62 */
    if (s->c[4] < 0.5f || s->c[6] < 0.5f || s->c[7] > 0.5f) return;
    int cand = candidate(s); int64_t P = P_(s), Ub = U_(s), Vb = V_(s), Wb = W_(s);
    WO_PARALLEL for (int64_t i = Ub; i < wo_min64(total, P); i++) {
      int axis = i < Vb ? 0 : (i < Wb ? 1 : 2); int64_t base = axis == 0 ? Ub : (axis == 1 ? Vb : Wb); I3 fn = n; fn.v[axis]++;
      I3 r = coord(i - base, fn), l = r; l.v[axis]--;
      float value = 0;
      if (!solid(s, l) && !solid(s, r) && (wet(s, l) || wet(s, r)))
        value = get_value(s, cand, i) - dt / 1000.0f * (press(s, r) - press(s, l)) / distance_face(s, l, r, axis);
      if (axis == 0 && r.v[0] == 0 && !solid(s, r)) value = get_value(s, cand, i);
      put_value(s, cand, i, value);
    }
    return;
  }
  if (op == 12) { /* pressure_pass.glslinc:63-77 */
    if (s->c[4] < 0.5f || s->c[6] < 0.5f || s->c[7] > 0.5f) return;
    int cand = candidate(s); float cell = PC_H(s, 0) * PC_H(s, 1) * PC_H(s, 2);
    FOR_GROUPS(s, groups) {
      F4 lane[128];
      for (int lid = 0; lid < 128; lid++) {
        int64_t i = g * 128 + lid; F4 value = f4(0, 0, 0, 0);
        if (i < N) {
          I3 q = coord(i, n); float speed = 0;
          for (int axis = 0; axis < 3; axis++) {
            I3 r = q; r.v[axis]++;
            speed = fmaxf_(speed, fmaxf_(fabsf(get_value(s, cand, face_index(s, q, axis))), fabsf(get_value(s, cand, face_index(s, r, axis)))));
          }
          float div = wet(s, q) ? dt * fabsf(divergence(s, q)) : 0;
          float err = wet(s, q) ? fabsf(press(s, q) - 9810.0f * (eta(s, q) - center(s, q).v[1])) : 0;
          value = f4(div, speed, err, get_value(s, cand, i) * PC_H(s, 0) * PC_H(s, 1) * PC_H(s, 2));
          (void)cell;
          int bad = 0; for (int k = 0; k < 4; k++) if (isnan(value.v[k]) || isinf(value.v[k])) bad = 1;
          if (bad) value = f4(1e30f, 1e30f, 1e30f, 1e30f);
        }
        lane[lid] = value;
      }
      s->partial[g] = tree_reduce(lane, RED_DIAG);
    }
    return;
  }
  if (op == 13) { /* pressure_pass.glslinc:78-103 */
    float *c = s->c;
    if (c[4] < 0.5f || c[6] < 0.5f || c[7] > 0.5f) return;
    F4 v = f4(0, 0, 0, 0); int64_t G = (N + 127) / 128;
    for (int64_t k = 0; k < G; k++) { for (int j = 0; j < 3; j++) v.v[j] = fmaxf_(v.v[j], s->partial[k].v[j]); v.v[3] += s->partial[k].v[3]; }
    c[14] = v.v[0]; c[15] = v.v[1]; c[16] = v.v[2]; c[17] = v.v[3];
    if ((v.v[0] >= 1e-4f || c[61] < .5f) && v.v[1] < 1e20f && c[19] < .5f && c[58] < 4) { c[57] = 1; c[58]++; c[6] = 0; return; }
    if (c[61] < .5f) { c[3]++; c[7] = 4; return; }
    if (v.v[0] >= 1e-4f || v.v[1] >= 1e20f || c[19] > 0.5f) { c[3]++; c[7] = 4; return; }
    c[28] = v.v[3] + c[40] + c[39] - c[26] - c[24] - c[22];
    c[28] += c[74] + c[66] - c[73] - c[65] - c[72] - c[64];
    if (fabsf(c[28]) > 1e-3f * fmaxf_(c[26], 1e-20f)) { c[3]++; c[7] = 21; return; }
    if (c[27] > .450001f) { c[3]++; c[7] = 20; return; }
    c[24] += c[22]; c[25] += c[23]; c[40] += c[39]; c[29] = 0; c[30] = 0;
    for (int field = 0; field < 6; field++) c[72 + field] += c[64 + field];
    c[90] = dt;
    c[0] = (float)candidate(s); c[1] += dt; c[2]++; c[4] = 0;
    return;
  }
  if (op == 14) { /* pressure_pass.glslinc:104-107 */
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(N, total); i++) {
      float f = get_value(s, accepted(s), i);
      s->img_planes[4 * i] = 0; s->img_planes[4 * i + 1] = 1; s->img_planes[4 * i + 2] = 0; s->img_planes[4 * i + 3] = f;
      s->img_kinds[i] = f == 0 ? 0 : (f == 1 ? 1 : 2);
    }
    return;
  }
  if (op == 16) { /* pressure_pass.glslinc:108-115 */
    if (s->c[4] > .5f && s->c[5] < .5f) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(N, total); i++) {
      I3 q = coord(i, n); float diag = 0; int face = 0; int w = wet(s, q);
      for (int axis = 0; axis < 3; axis++) for (int sign = -1; sign <= 1; sign += 2) {
        I3 nb = q; nb.v[axis] += sign; float k = w ? coefficient(s, q, nb, axis) : 0;
        s->work[(int64_t)(6 + face) * N + i] = k; diag += k; face++;
      }
      s->work[12 * N + i] = w ? diag : 1.0f;
    }
    return;
  }
  if (op == 17) { /* pressure_pass.glslinc:116-126 */
    if (s->c[4] > .5f && s->c[5] < .5f) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(N, total); i++) { F4 e = wo_reconstruct_pressure(s, coord(i, n)); memcpy(s->plane + 4 * i, e.v, 16); }
    return;
  }
  if (op == 19) { /* pressure_pass.glslinc:127-130 */
    if (s->c[4] > .5f) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(N, total); i++) {
      I3 q = coord(i, n); float f = get_value(s, accepted(s), i);
      F4 e = wo_reconstruct_pressure(s, q); memcpy(s->img_planes + 4 * i, e.v, 16);
      s->img_kinds[i] = f <= 0 ? 0 : (f >= 1 ? 1 : 2);
    }
    return;
  }
  if (op == 18) { /* pressure_pass.glslinc:131-137 */
    if (s->c[4] > .5f && s->c[5] < .5f) return;
    WO_PARALLEL for (int64_t i = 0; i < wo_min64(N, total); i++) {
      I3 q = coord(i, n); float ph = wo_geometry_distance(s, q);
      s->geo[2 * i] = ph; s->geo[2 * i + 1] = solid(s, q) ? -1.0f : (ph < 0 ? 1.0f : 0.0f);
    }
    return;
  }
  if (op == 15) { s->c[19] = 1; return; }
}
