/* Internal state and per-invocation helpers of the oracle (test infrastructure).
 * Each helper names the reference function it restates. */
#ifndef WO_STATE_H
#define WO_STATE_H
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "wo.h"

typedef struct { int v[3]; } I3;
typedef struct { float v[3]; } F3;
typedef struct { float v[4]; } F4;

#define NCTRL 96

/* The invocations of one op write disjoint outputs (as on the GPU), so an
 * op's invocation loop may run on several host threads (OpenMP); the
 * per-group reduction partials and the serial finalisers keep their order,
 * so the results do not depend on the thread count.  Ops whose serial loop
 * accumulates into shared state (the multigrid scatter) stay serial. */
#ifdef _OPENMP
#define WO_PARALLEL _Pragma("omp parallel for schedule(dynamic, 512)")
#define WO_PARALLEL_COUNT(x) _Pragma(WO_STR(omp parallel for schedule(static) reduction(+:x)))
#define WO_STR(x) #x
#else
#define WO_PARALLEL
#define WO_PARALLEL_COUNT(x)
#endif
static inline int64_t wo_min64(int64_t a, int64_t b) { return a < b ? a : b; }
#define MAXL 12

struct WO {
  float pc[32];
  float opts[8];
  int op;                 /* current dispatch op (geometry_bank reads it) */
  int nx, ny, nz;
  int64_t N, NU, NV, NW, NF, stride;
  float *a, *b, *work, *geo, *c, *plane, *motion, *history, *chist;
  float probe[32];
  float *img_planes, *img_kinds;
  F4 *partial; int64_t npartial;
  float boundary_options[4];
  /* multigrid (water/pressure_multigrid.gd) */
  int L;
  int *meta; int nmeta;
  uint32_t *matrix; int64_t matrix_count;
  float *vec; int64_t node_count;
  float *rows; int64_t row_count;       /* vec2 pairs */
  float *transfer; int64_t transfer_count; /* vec2 pairs */
  float *snapshot, *previous; uint32_t *state; int64_t state_count;
  int smoothing, lag, aggregate_from, mode, iters_per_batch, history_capacity;
  int (*cycle)[4]; int ncycle;
  int (*build)[4]; int nbuild;
  int mg_phase;            /* run.z of the current MG dispatch */
  int mg_run[4];
  int scatter_reverse;     /* Galerkin scatter (MG op 3) in descending entry order */
};

static inline I3 i3(int x, int y, int z) { I3 r; r.v[0] = x; r.v[1] = y; r.v[2] = z; return r; }
static inline F3 f3(float x, float y, float z) { F3 r; r.v[0] = x; r.v[1] = y; r.v[2] = z; return r; }
static inline F4 f4(float x, float y, float z, float w) { F4 r; r.v[0] = x; r.v[1] = y; r.v[2] = z; r.v[3] = w; return r; }
static inline float dot3(F3 a, F3 b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
static inline float len3(F3 a) { return sqrtf(dot3(a, a)); }
static inline float fminf_(float a, float b) { return b < a ? b : a; }  /* GLSL min */
static inline float fmaxf_(float a, float b) { return a < b ? b : a; }  /* GLSL max */
static inline float clampf(float x, float lo, float hi) { return fminf_(fmaxf_(x, lo), hi); }
static inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
static inline float fractf(float x) { return x - floorf(x); }
static inline float signf(float x) { return x > 0 ? 1.0f : (x < 0 ? -1.0f : 0.0f); }

/* push-constant views, pressure_core.glslinc:13-22 */
#define PC_H(s, ax) ((s)->pc[4 + (ax)])
#define PC_DTMAX(s) ((s)->pc[7])
#define PC_O(s, ax) ((s)->pc[8 + (ax)])
#define PC_ETA(s) ((s)->pc[11])
#define PC_PERX(s) ((s)->pc[12])
#define PC_SLOPE(s) ((s)->pc[13])
#define PC_CORR(s) ((s)->pc[14])
#define PC_QUOTA(s) ((s)->pc[15])
#define PC_FIX(s, k) ((s)->pc[16 + (k)])
#define PC_WAVE(s, k) ((s)->pc[20 + (k)])
#define PC_NUM(s, k) ((s)->pc[24 + (k)])
#define PC_FORC(s, k) ((s)->pc[28 + (k)])

/* periodic_index.glslinc:3-6 */
static inline int pidx(int q, int n) { if (q >= 0) return q % n; return n - 1 - ((-1 - q) % n); }

static inline I3 dims(const WO *s) { return i3(s->nx, s->ny, s->nz); }
static inline int64_t U_(const WO *s) { return s->N; }
static inline int64_t V_(const WO *s) { return s->N + s->NU; }
static inline int64_t W_(const WO *s) { return s->N + s->NU + s->NV; }
static inline int64_t P_(const WO *s) { return s->N + s->NU + s->NV + s->NW; }
static inline int64_t index3(I3 q, I3 n) { return (int64_t)q.v[0] + (int64_t)n.v[0] * (q.v[1] + (int64_t)n.v[1] * q.v[2]); }
static inline I3 coord(int64_t i, I3 n) { return i3((int)(i % n.v[0]), (int)((i / n.v[0]) % n.v[1]), (int)(i / ((int64_t)n.v[0] * n.v[1]))); }
static inline float get_value(const WO *s, int bank, int64_t i) { return bank == 0 ? s->a[i] : s->b[i]; }
static inline void put_value(WO *s, int bank, int64_t i, float v) { if (bank == 0) s->a[i] = v; else s->b[i] = v; }
static inline float timestep(const WO *s) { return s->c[20]; }
static inline int accepted(const WO *s) { return (int)s->c[0]; }
static inline int candidate(const WO *s) { return 1 - accepted(s); }
static inline int geometry_bank(const WO *s) { return (s->op == 19 || s->c[4] < .5f) ? accepted(s) : candidate(s); }

/* pressure_core.glslinc:73-83 */
static inline float eta(const WO *s, I3 q) {
  if (PC_FIX(s, 0) == 5) {
    float phase = PC_WAVE(s, 1) * (PC_O(s, 0) + ((float)q.v[0] + .5f) * PC_H(s, 0));
    float sigma = tanhf(PC_WAVE(s, 1) * PC_WAVE(s, 3));
    float second = PC_NUM(s, 2) > .5f ? PC_WAVE(s, 1) * PC_WAVE(s, 0) * PC_WAVE(s, 0) * (3 - sigma * sigma) / (4 * sigma * sigma * sigma) : 0;
    return PC_ETA(s) + PC_WAVE(s, 0) * cosf(phase) + second * cosf(2 * phase);
  }
  return PC_ETA(s) + PC_CORR(s) * cosf(6.28318530718f * ((float)q.v[0] + 0.5f) / (float)s->nx);
}
static inline F3 center(const WO *s, I3 q) {
  F3 r; for (int k = 0; k < 3; k++) r.v[k] = PC_O(s, k) + ((float)q.v[k] + 0.5f) * PC_H(s, k); return r;
}
static inline I3 wrap_cell(const WO *s, I3 q) {
  q.v[2] = pidx(q.v[2], s->nz); if (PC_PERX(s) > 0.5f) q.v[0] = pidx(q.v[0], s->nx); return q;
}
static inline int outside(const WO *s, I3 q) { return q.v[0] < 0 || q.v[0] >= s->nx || q.v[1] < 0 || q.v[1] >= s->ny; }
/* pressure_core.glslinc:87-96 (COASTAL: x<0 is open ocean) */
static inline int solid(const WO *s, I3 q) {
  q = wrap_cell(s, q);
  if (q.v[1] < 0 || q.v[0] >= s->nx) return 1;
  if (q.v[1] >= s->ny) return 0;
  if (PC_SLOPE(s) > 0.0f) {
    float x = center(s, q).v[0];
    float bed = -0.45f + fmaxf_(0.0f, x - 1.5f) / PC_SLOPE(s);
    return center(s, q).v[1] <= bed;
  }
  return 0;
}
static inline int wet(const WO *s, I3 q) { q = wrap_cell(s, q); return !outside(s, q) && s->geo[2 * index3(q, dims(s)) + 1] > 0.5f; }
static inline float phi(const WO *s, I3 q) { q = wrap_cell(s, q); return outside(s, q) ? center(s, q).v[1] - eta(s, q) : s->geo[2 * index3(q, dims(s))]; }
static inline float press(const WO *s, I3 q) { q = wrap_cell(s, q); return wet(s, q) ? get_value(s, candidate(s), P_(s) + index3(q, dims(s))) : 0.0f; }
/* pressure_core.glslinc:100-109 */
static inline float distance_face(const WO *s, I3 l, I3 r, int axis) {
  float h = PC_H(s, axis);
  if (axis == 1 && (l.v[1] >= s->ny || r.v[1] >= s->ny)) return .5f * h;
  int wl = wet(s, l), wr = wet(s, r);
  if (wl && wr) return h;
  float pl = fabsf(phi(s, l)), pr = fabsf(phi(s, r));
  return h * fmaxf_((wl ? pl : pr) / fmaxf_(pl + pr, 1e-30f), 1e-6f);
}
/* pressure_core.glslinc:110-118 */
static inline float coefficient(const WO *s, I3 q, I3 nb, int axis) {
  if (solid(s, nb)) return 0.0f;
  if (axis == 0 && nb.v[0] < 0) return 0.0f;
  return 1.0f / (PC_H(s, axis) * distance_face(s, q, nb, axis));
}
static inline float vector_at(const WO *s, I3 q, int which) {
  q = wrap_cell(s, q); if (!wet(s, q)) return 0.0f;
  int64_t i = index3(q, dims(s));
  return which == 0 ? press(s, q) : s->work[2 * s->N + i];
}
/* pressure_core.glslinc:120-125 */
static inline float operator_at(const WO *s, I3 q, int which) {
  float result = 0, here = vector_at(s, q, which); int64_t i = index3(q, dims(s));
  int face = 0;
  for (int axis = 0; axis < 3; axis++) for (int sign = -1; sign <= 1; sign += 2) {
    I3 nb = q; nb.v[axis] += sign; float k = s->work[(int64_t)(6 + face) * s->N + i];
    result += k * (here - vector_at(s, nb, which)); face++;
  }
  return result;
}
static inline float diagonal(const WO *s, I3 q) { return s->work[12 * s->N + index3(q, dims(s))]; }
/* pressure_core.glslinc:129-133 */
static inline int64_t face_index(const WO *s, I3 q, int axis) {
  I3 n = dims(s); if (axis == 2 && q.v[2] == n.v[2]) q.v[2] = 0;
  if (axis == 0 && PC_PERX(s) > .5f && q.v[0] == n.v[0]) q.v[0] = 0;
  n.v[axis]++;
  return (axis == 0 ? U_(s) : (axis == 1 ? V_(s) : W_(s))) + index3(q, n);
}
static inline float divergence(const WO *s, I3 q) {
  float d = 0;
  for (int axis = 0; axis < 3; axis++) {
    I3 r = q; r.v[axis]++;
    d += (get_value(s, candidate(s), face_index(s, r, axis)) - get_value(s, candidate(s), face_index(s, q, axis))) / PC_H(s, axis);
  }
  return d;
}
/* pressure_core.glslinc:156-179 */
static inline int pressure_setup(const WO *s) { return s->c[5] > .5f || s->c[57] == 2; }
static inline float pressure_scaled_norm(const WO *s, float sq) { return timestep(s) * timestep(s) / 1000 * sqrtf(fmaxf_(sq, 0)); }
static inline int pressure_norm_converged(const WO *s, float r2, float b2) {
  float tol = PC_NUM(s, 0);
  return r2 <= tol * tol * b2 || (pressure_scaled_norm(s, b2) <= 1e-5f && pressure_scaled_norm(s, r2) <= 1e-8f);
}
static inline int residual_converged(const WO *s, F4 sums, float b2) {
  return pressure_norm_converged(s, sums.v[1], b2) && timestep(s) * timestep(s) / 1000 * sums.v[3] < 5e-5f;
}
static inline int solving(const WO *s) { return s->c[4] > 0.5f && s->c[6] < 0.5f && s->c[7] < 0.5f; }

/* motion layout, coupled_motion.glslinc:8-14 */
static inline int64_t faces_(const WO *s) { return s->NU + s->NV + s->NW; }
static inline int64_t flux_base(const WO *s) { return 2 * s->N; }
static inline int64_t forward_base(const WO *s) { return flux_base(s) + faces_(s); }
static inline int64_t source_base(const WO *s, int bank) { return forward_base(s) + faces_(s) + bank * faces_(s); }
static inline int64_t valid_base(const WO *s, int bank) { return source_base(s, 0) + 2 * faces_(s) + bank * faces_(s); }
static inline int64_t pressure_mask_base(const WO *s, int bank) { return valid_base(s, 0) + 2 * faces_(s) + bank * s->N; }
static inline int64_t trace_cache_base(const WO *s) { return 4 * s->N + 6 * faces_(s); }
static inline int64_t face_local(const WO *s, I3 q, int axis) { return face_index(s, q, axis) - U_(s); }
static inline void unpack_face(const WO *s, int64_t f, I3 *q, int *axis) {
  *axis = f < s->NU ? 0 : (f < s->NU + s->NV ? 1 : 2);
  int64_t base = *axis == 0 ? 0 : (*axis == 1 ? s->NU : s->NU + s->NV);
  I3 n = dims(s); n.v[*axis]++; *q = coord(f - base, n);
}

/* forward declarations of the larger helpers (wo_geometry.c) */
float wo_cube_volume3(F3 normal, float alpha);
float wo_inverse_volume3(F3 normal, float fraction);
float wo_geometry_fraction(const WO *s, I3 q);
float wo_transport_f(const WO *s, I3 q);
F3 wo_interface_gradient(const WO *s, I3 q, int source);
F3 wo_height_normal(const WO *s, I3 q, F3 youngs, int source, int *used);
F4 wo_reconstruct_pressure(const WO *s, I3 q);
F4 wo_transport_plane(const WO *s, I3 q);
float wo_geometry_distance(const WO *s, I3 q);
/* coastal_target.glslinc */
float wo_smooth_ramp(float q);
void wo_coastal_surface(const WO *s, float x, float time, float out2[2]);
F3 wo_coastal_velocity(const WO *s, F3 world, float time);
F3 wo_coastal_face_position(const WO *s, I3 q, int axis);
F4 wo_coastal_plane(const WO *s, I3 q, float time);
float wo_coastal_fraction(const WO *s, I3 q, float time);
float wo_coastal_rate(const WO *s, float x);
/* op families */
void wo_pressure_op(WO *s, int op, int64_t groups);
void wo_motion_op(WO *s, int op, int64_t groups);
void wo_coastal_op(WO *s, int op, int64_t groups);
void wo_mg_op(WO *s, int op, int level, int phase, int field, int64_t groups);
void wo_audit_op(WO *s, int op, int64_t groups);
int wo_mg_init(WO *s);
void wo_mg_free(WO *s);

/* 128-lane workgroup tree reductions (pressure_core.glslinc:135-155,
 * coupled_motion.glslinc:57-61, pressure_pass.glslinc:74-76). */
enum { RED_SUM = 0, RED_PRESSURE = 1, RED_BOUNDS = 2, RED_DIAG = 3 };
static inline F4 red_combine(F4 a, F4 b, int mode) {
  F4 r = a;
  switch (mode) {
  case RED_SUM: for (int k = 0; k < 4; k++) r.v[k] = a.v[k] + b.v[k]; break;
  case RED_PRESSURE: for (int k = 0; k < 3; k++) r.v[k] = a.v[k] + b.v[k]; r.v[3] = fmaxf_(a.v[3], b.v[3]); break;
  case RED_BOUNDS: r.v[0] = fminf_(a.v[0], b.v[0]); r.v[1] = fmaxf_(a.v[1], b.v[1]); r.v[2] = a.v[2] + b.v[2]; r.v[3] = a.v[3] + b.v[3]; break;
  case RED_DIAG: for (int k = 0; k < 3; k++) r.v[k] = fmaxf_(a.v[k], b.v[k]); r.v[3] = a.v[3] + b.v[3]; break;
  }
  return r;
}
static inline F4 tree_reduce(F4 *lane, int mode) {
  for (int off = 64; off > 0; off >>= 1) for (int lid = 0; lid < off; lid++) lane[lid] = red_combine(lane[lid], lane[lid + off], mode);
  return lane[0];
}
/* global_reduce(max_w), pressure_core.glslinc:146-155 */
static inline F4 global_reduce(const WO *s, int max_w) {
  F4 lane[128]; int64_t G = (s->N + 127) / 128;
  for (int lid = 0; lid < 128; lid++) {
    F4 v = f4(0, 0, 0, 0);
    for (int64_t k = lid; k < G; k += 128) {
      for (int j = 0; j < 3; j++) v.v[j] += s->partial[k].v[j];
      v.v[3] = max_w ? fmaxf_(v.v[3], s->partial[k].v[3]) : v.v[3] + s->partial[k].v[3];
    }
    lane[lid] = v;
  }
  return tree_reduce(lane, max_w ? RED_PRESSURE : RED_SUM);
}
#endif
