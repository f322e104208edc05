/* Oracle lifecycle and host sequencing (test infrastructure). Restates the
 * buffer allocation of water/pressure_gpu.gd:82-127 and water/coupled_gpu.gd:
 * 23-33, and the command order of water/coupled_gpu.gd:34-56 and
 * water/coastal_gpu.gd:26-33. */
#include <stdio.h>
#include "wo_state.h"

WO *wo_create(const float *pc, const float *opts) {
  WO *s = calloc(1, sizeof(WO));
  memcpy(s->pc, pc, sizeof s->pc);
  memcpy(s->opts, opts, sizeof s->opts);
  s->nx = (int)pc[0]; s->ny = (int)pc[1]; s->nz = (int)pc[2];
  if (s->nx < 1 || s->ny < 1 || s->nz < 1) { free(s); return NULL; }
  s->N = (int64_t)s->nx * s->ny * s->nz;
  s->NU = (int64_t)(s->nx + 1) * s->ny * s->nz; s->NV = (int64_t)s->nx * (s->ny + 1) * s->nz; s->NW = (int64_t)s->nx * s->ny * (s->nz + 1);
  s->NF = s->NU + s->NV + s->NW; s->stride = 2 * s->N + s->NF;
  s->boundary_options[0] = opts[0];
  s->smoothing = (int)opts[1]; s->lag = (int)opts[2]; s->aggregate_from = (int)opts[3]; s->mode = (int)opts[4]; s->iters_per_batch = (int)opts[5];
  s->history_capacity = (int)pc[19];
  s->a = calloc((size_t)s->stride, 4); s->b = calloc((size_t)s->stride, 4);
  s->work = calloc((size_t)s->N * 14, 4); s->geo = calloc((size_t)s->N * 2, 4);
  s->c = calloc(NCTRL, 4); s->plane = calloc((size_t)s->N * 4, 4);
  s->motion = calloc((size_t)(4 * s->N + 12 * s->NF), 4);
  s->history = calloc((size_t)s->history_capacity * 16, 4);
  s->chist = calloc((size_t)s->history_capacity * 20, 4);
  s->img_planes = calloc((size_t)s->N * 4, 4); s->img_kinds = calloc((size_t)s->N, 4);
  s->npartial = (s->stride + 127) / 128 + 1;
  s->partial = calloc((size_t)s->npartial, sizeof(F4));
  if (wo_mg_init(s) != 0) { wo_destroy(s); return NULL; }
  return s;
}

void wo_destroy(WO *s) {
  if (!s) return;
  free(s->a); free(s->b); free(s->work); free(s->geo); free(s->c); free(s->plane); free(s->motion);
  free(s->history); free(s->chist); free(s->img_planes); free(s->img_kinds); free(s->partial);
  wo_mg_free(s); free(s);
}

static void *field_ptr(WO *s, int field, int64_t *count, int *bytes) {
  *bytes = 4;
  switch (field) {
  case 0: *count = s->stride; return s->a;
  case 1: *count = s->stride; return s->b;
  case 2: *count = 14 * s->N; return s->work;
  case 3: *count = 2 * s->N; return s->geo;
  case 4: *count = NCTRL; return s->c;
  case 5: *count = 4 * s->N; return s->plane;
  case 6: *count = 4 * s->N + 12 * s->NF; return s->motion;
  case 7: *count = 4 * s->N; return s->img_planes;
  case 8: *count = s->N; return s->img_kinds;
  case 9: *count = (int64_t)s->history_capacity * 16; return s->history;
  case 10: *count = (int64_t)s->history_capacity * 20; return s->chist;
  case 11: *count = 32; return s->probe;
  case 12: *count = 4 * ((s->N + 127) / 128); return s->partial;
  case 13: *count = 8 * s->node_count; return s->vec;
  case 14: *count = 9 * s->N; return s->snapshot;
  case 15: *count = s->state_count; return s->state;
  case 16: *count = 2 * s->row_count; return s->rows;
  case 17: { int64_t n = s->meta[16 * (s->L - 1) + 8]; *count = n * n; return s->matrix + s->meta[16 * (s->L - 1) + 9]; }
  case 19: *count = 9 * s->N; return s->previous;
  }
  *count = 0; return NULL;
}
int64_t wo_field_count(WO *s, int field) { int64_t n; int b; field_ptr(s, field, &n, &b); return n; }
int wo_get(WO *s, int field, void *dst, int64_t count) {
  int64_t n; int b; void *p = field_ptr(s, field, &n, &b); if (!p || count > n) return -1;
  memcpy(dst, p, (size_t)count * b); return 0;
}
int wo_set(WO *s, int field, const void *src, int64_t count) {
  int64_t n; int b; void *p = field_ptr(s, field, &n, &b); if (!p || count > n) return -1;
  memcpy(p, src, (size_t)count * b); return 0;
}

void wo_set_quota(WO *s, float quota) { s->pc[15] = quota; }
void wo_set_scatter_order(WO *s, int reverse) { s->scatter_reverse = reverse != 0; }

/* coastal.glsl:15 routing; op 90 is the benchmark perturbation (no reference op) */
void wo_op(WO *s, int op, int64_t groups) {
  s->pc[3] = (float)op; s->op = op;
  if (op == 90) wo_pressure_op(s, op, groups);
  else if (op >= 80) wo_coastal_op(s, op, groups);
  else if (op >= 20) wo_motion_op(s, op, groups);
  else wo_pressure_op(s, op, groups);
}

static int64_t cg_(WO *s) { return (s->N + 127) / 128; }
static int64_t sg_(WO *s) { return (s->stride + 127) / 128; }
static int64_t fg_(WO *s) { return (s->NF + 127) / 128; }

/* water/coupled_gpu.gd:34-37, water/coastal_gpu.gd:26-27 */
void wo_initial(WO *s) {
  s->pc[15] = 0;
  wo_op(s, 0, sg_(s));
  wo_op(s, 40, sg_(s));
  if (s->opts[6] != 0) wo_op(s, 90, cg_(s));
  int ops[] = {17, 18, 16, 19, 41, 42, 49};
  for (int k = 0; k < 7; k++) wo_op(s, ops[k], ops[k] == 42 ? 1 : (ops[k] == 49 ? sg_(s) : cg_(s)));
  wo_op(s, 85, 1); wo_op(s, 86, 1);
}

/* one batch: water/coupled_gpu.gd:38-56 with water/coastal_gpu.gd:28-33, and
 * the MG hooks of water/pressure_gpu.gd:69-74 */
static void dispatch(WO *s, int op, int64_t groups) {
  if (op == 200) { wo_mg_record(s, 1, 0); return; }
  if (op == 4 || op == 9) wo_mg_record(s, 0, op == 4);
  wo_op(s, op, groups);
}
void wo_batch(WO *s, float quota) {
  s->pc[15] = quota;
  int64_t cg = cg_(s), sg = sg_(s), fg = fg_(s);
  dispatch(s, 43, 1); dispatch(s, 1, 1);
  dispatch(s, 20, cg); dispatch(s, 21, 1); dispatch(s, 45, cg); dispatch(s, 46, 1); dispatch(s, 22, sg);
  for (int axis = 0; axis < 3; axis++) {
    dispatch(s, 23, cg); dispatch(s, 24, fg); dispatch(s, 55, cg); dispatch(s, 56, 1); dispatch(s, 25, cg); dispatch(s, 26, 1);
  }
  dispatch(s, 27, fg); dispatch(s, 28, fg);
  { int ops[] = {80, 81, 82, 83, 84, 87, 88}; for (int k = 0; k < 7; k++) dispatch(s, ops[k], (ops[k] == 81 || ops[k] == 84 || ops[k] == 88) ? 1 : (ops[k] == 82 ? fg : cg)); }
  dispatch(s, 17, cg); dispatch(s, 18, cg); dispatch(s, 16, cg);
  dispatch(s, 2, sg); dispatch(s, 47, fg); dispatch(s, 63, 1); dispatch(s, 64, cg);
  dispatch(s, 200, cg);
  dispatch(s, 3, cg); dispatch(s, 4, 1);
  for (int it = 0; it < s->iters_per_batch; it++) { dispatch(s, 5, cg); dispatch(s, 6, cg); dispatch(s, 7, 1); dispatch(s, 8, cg); dispatch(s, 9, 1); dispatch(s, 10, cg); }
  dispatch(s, 11, sg); dispatch(s, 65, cg); dispatch(s, 47, fg);
  { int ops[] = {30, 31, 32, 33, 34, 47, 48}; for (int k = 0; k < 7; k++) dispatch(s, ops[k], fg); }
  dispatch(s, 35, cg); dispatch(s, 36, 1);
  dispatch(s, 62, cg); dispatch(s, 66, 1); dispatch(s, 12, cg); dispatch(s, 13, 1); dispatch(s, 44, 1); dispatch(s, 19, cg);
  dispatch(s, 85, 1); dispatch(s, 89, 1);
}
