/* Multigrid preconditioner and hierarchy audit of the oracle (test
 * infrastructure). Restates water/pressure_multigrid.gd (level planning and
 * command lists), water/shaders/compute/pressure_multigrid.glslinc,
 * pressure_mg_coarse.glslinc and pressure_hierarchy_state.glsl. */
#include <stdio.h>
#include "wo_state.h"

static int md(const WO *s, int l, int k) { return s->meta[16 * l + k]; }
static I3 mdims(const WO *s, int l) { return i3(md(s, l, 0), md(s, l, 1), md(s, l, 2)); }
static int mcount(const WO *s, int l) { return md(s, l, 8); }
static int64_t at(const WO *s, int l, int field, int64_t i) { return 8 * (int64_t)md(s, l, 3) + (int64_t)field * mcount(s, l) + i; }
static int pmod(int x, int n) { return x >= 0 ? x % n : (n - 1) - ((-1 - x) % n); }
static int perx(const WO *s) { return PC_PERX(s) > .5f; }
/* pressure_multigrid.glslinc:36-41 */
static int mwrap(const WO *s, int l, I3 q) {
  I3 n = mdims(s, l); q.v[2] = pmod(q.v[2], n.v[2]);
  if (perx(s)) q.v[0] = pmod(q.v[0], n.v[0]);
  if (q.v[0] < 0 || q.v[0] >= n.v[0] || q.v[1] < 0 || q.v[1] >= n.v[1]) return -1;
  return (int)index3(q, n);
}
static I3 mdelta(const WO *s, int l, int i, int j) {
  I3 n = mdims(s, l), a = coord(i, n), b = coord(j, n), d;
  for (int k = 0; k < 3; k++) d.v[k] = b.v[k] - a.v[k];
  if (perx(s)) d.v[0] = pmod(d.v[0] + n.v[0] / 2, n.v[0]) - n.v[0] / 2;
  d.v[2] = pmod(d.v[2] + n.v[2] / 2, n.v[2]) - n.v[2] / 2;
  return d;
}
static int slot(I3 d) { return d.v[0] + 2 + 5 * (d.v[1] + 2 + 5 * (d.v[2] + 2)); }
static I3 moffset(const WO *s, int l, int k) { int st = md(s, l, 10) + 3 * k; return i3(s->meta[st], s->meta[st + 1], s->meta[st + 2]); }
static int active_cell(const WO *s, int l, int i) { return i >= 0 && s->vec[at(s, l, 3, i)] > .5f; }
static float fine_coefficient(const WO *s, int field, int i) { return s->snapshot[(int64_t)(field - 6) * mcount(s, 0) + i]; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float i2f(int i) { float f; memcpy(&f, &i, 4); return f; }
static int f2i(float f) { int i; memcpy(&i, &f, 4); return i; }
static int any_gt2(I3 d) { return abs(d.v[0]) > 2 || abs(d.v[1]) > 2 || abs(d.v[2]) > 2; }
/* pressure_multigrid.glslinc:74-88 (level >= 1) */
static float mcoefficient(const WO *s, int l, int i, int j) {
  if (j < 0) return 0;
  if (!active_cell(s, l, i) || !active_cell(s, l, j)) return i == j ? 1 : 0;
  int a = i < j ? i : j, b = i < j ? j : i; I3 d = mdelta(s, l, a, b);
  if (any_gt2(d)) return 0;
  return u2f(s->matrix[md(s, l, 4) + 125 * (int64_t)a + slot(d)]);
}
static int unique_neighbor(const WO *s, int l, int i, I3 off, int j) {
  if (j < 0) return 0; I3 d = mdelta(s, l, i, j);
  return d.v[0] == off.v[0] && d.v[1] == off.v[1] && d.v[2] == off.v[2];
}
/* pressure_multigrid.glslinc:90-105 */
static float apply(const WO *s, int l, int field, int i) {
  if (!active_cell(s, l, i)) return s->vec[at(s, l, field, i)];
  float sum = 0; I3 q = coord(i, mdims(s, l));
  if (l == 0) {
    float here = s->vec[at(s, l, field, i)];
    for (int axis = 0; axis < 3; axis++) for (int side = 0; side < 2; side++) {
      I3 nb = q; nb.v[axis] += 2 * side - 1; int j = mwrap(s, l, nb);
      float other = active_cell(s, l, j) ? s->vec[at(s, l, field, j)] : 0;
      sum += fine_coefficient(s, 6 + 2 * axis + side, i) * (here - other);
    }
    return sum;
  }
  int width = md(s, l, 11), N = mcount(s, l); int64_t start = md(s, l, 12) + (int64_t)i;
  for (int k = 0; k < width; k++) { const float *e = s->rows + 2 * (start + (int64_t)k * N); sum += e[0] * s->vec[at(s, l, field, f2i(e[1]))]; }
  return sum;
}
/* pressure_multigrid.glslinc:108-120 */
static int compute_transfers(const WO *s, int l, int i, int *ids, float *w) {
  I3 n = mdims(s, l + 1), f = i3(md(s, l, 5), md(s, l, 6), md(s, l, 7)), q = coord(i, mdims(s, l));
  if (md(s, l, 15) != 0) {
    I3 qc; for (int k = 0; k < 3; k++) qc.v[k] = q.v[k] / f.v[k];
    int j = (int)index3(qc, n); if (!active_cell(s, l + 1, j)) return 0; ids[0] = j; w[0] = 1; return 1;
  }
  I3 base; F3 t;
  for (int k = 0; k < 3; k++) { float pos = ((float)q.v[k] + .5f) / (float)f.v[k] - .5f; base.v[k] = (int)floorf(pos); t.v[k] = fractf(pos); }
  int size = 0;
  for (int z = 0; z < f.v[2]; z++) for (int y = 0; y < f.v[1]; y++) for (int x = 0; x < f.v[0]; x++) {
    int bit[3] = {x, y, z}; I3 nb = i3(base.v[0] + x, base.v[1] + y, base.v[2] + z); float ww = 1;
    for (int axis = 0; axis < 3; axis++) if (f.v[axis] > 1) ww *= bit[axis] == 1 ? t.v[axis] : 1 - t.v[axis];
    if (nb.v[1] >= n.v[1]) continue; if (nb.v[1] < 0) nb.v[1] = 0;
    nb.v[0] = perx(s) ? pmod(nb.v[0], n.v[0]) : clampi(nb.v[0], 0, n.v[0] - 1); nb.v[2] = pmod(nb.v[2], n.v[2]);
    int j = (int)index3(nb, n); if (!active_cell(s, l + 1, j) || ww == 0) continue;
    ids[size] = j; w[size] = ww; size++;
  }
  return size;
}
static int transfers(const WO *s, int l, int i, int *ids, float *w) {
  int size = 0;
  for (int k = 0; k < md(s, l, 14); k++) {
    const float *e = s->transfer + 2 * ((int64_t)md(s, l, 13) + (int64_t)k * mcount(s, l) + i);
    if (e[1] > 0) { ids[size] = f2i(e[0]); w[size] = e[1]; size++; }
  }
  return size;
}
static void add_entry(WO *s, int l, int i, int j, float value) {
  if (value == 0) return;
  int a = i < j ? i : j, b = i < j ? j : i; I3 d = mdelta(s, l, a, b);
  if (any_gt2(d)) { s->c[7] = 31; return; }
  int64_t address = md(s, l, 4) + 125 * (int64_t)a + slot(d);
  s->matrix[address] = f2u(u2f(s->matrix[address]) + value);
}
/* pressure_multigrid.glslinc:50-65 and pressure_mg_coarse.glslinc:27-34 */
static int mg_enabled(const WO *s, int op, int l) {
  const float *c = s->c;
  if (c[7] > .5f) return 0;
  if (s->mg_phase == 2) {
    if (c[5] < .5f) return 0;
    if (op == 1) return s->state[9] != 0;
    if (op == 16) return s->state[14] == 0 || s->state[20 + l + 1] != 0;
    return s->state[8] != 0;
  }
  if (s->mg_phase == 0) return (c[5] > .5f || c[57] == 2) && c[6] < .5f;
  return c[4] > .5f && c[6] < .5f;
}

static void coarse_op(WO *s, int op, int l) {
  int N = mcount(s, l);
  if (op == 11) { /* pressure_mg_coarse.glslinc:35-49 */
    float *dense = calloc((size_t)N * N, 4), *inv = calloc((size_t)N * N, 4);
    for (int k = 0; k < N * N; k++) dense[k] = mcoefficient(s, l, k / N, k % N);
    for (int k = 0; k < N; k++) {
      float d = dense[k * N + k]; for (int j = 0; j < k; j++) d -= dense[k * N + j] * dense[k * N + j];
      if (!(d > 0) || isnan(d) || isinf(d)) { s->c[7] = 33; d = 1; }
      dense[k * N + k] = sqrtf(d);
      for (int lid = k + 1; lid < N; lid++) { float e = dense[lid * N + k]; for (int j = 0; j < k; j++) e -= dense[lid * N + j] * dense[k * N + j]; dense[lid * N + k] = e / dense[k * N + k]; }
    }
    for (int row = 0; row < N; row++) for (int lid = 0; lid < N; lid++) {
      float d = lid == row ? 1 : 0; for (int j = 0; j < row; j++) d -= dense[row * N + j] * inv[j * N + lid];
      inv[row * N + lid] = d / dense[row * N + row];
    }
    for (int k = 0; k < N * N; k++) s->matrix[md(s, l, 9) + k] = f2u(inv[k]);
    free(dense); free(inv); return;
  }
  if (op == 12) { /* pressure_mg_coarse.glslinc:50-53 */
    float rhs[64]; const uint32_t *Li = s->matrix + md(s, l, 9);
    for (int lid = 0; lid < N; lid++) { float d = 0; for (int j = 0; j <= lid; j++) d += u2f(Li[lid * N + j]) * s->vec[at(s, l, 0, j)]; rhs[lid] = d; }
    for (int lid = 0; lid < N; lid++) { float d = 0; for (int j = lid; j < N; j++) d += u2f(Li[j * N + lid]) * rhs[j]; s->vec[at(s, l, 1, lid)] = d; }
  }
}

void wo_mg_op(WO *s, int op, int l, int phase, int field, int64_t groups) {
  (void)groups;
  s->mg_phase = phase;
  if (!mg_enabled(s, op, l)) return;
  if (op == 11 || op == 12) { coarse_op(s, op, l); return; }
  int N = mcount(s, l); I3 nl = mdims(s, l);
  switch (op) {
  case 1: { /* pressure_multigrid.glslinc:139-150 */
    uint32_t changed = 0;
    WO_PARALLEL_COUNT(changed)
    for (int i = 0; i < N; i++) {
      float value = 0;
      if (l == 0) value = s->geo[2 * i + 1] > .5f ? 1 : 0;
      else {
        I3 f = i3(md(s, l - 1, 5), md(s, l - 1, 6), md(s, l - 1, 7)), base = coord(i, nl);
        for (int k = 0; k < 3; k++) base.v[k] *= f.v[k];
        for (int z = 0; z < f.v[2]; z++) for (int y = 0; y < f.v[1]; y++) for (int x = 0; x < f.v[0]; x++)
          if (active_cell(s, l - 1, mwrap(s, l - 1, i3(base.v[0] + x, base.v[1] + y, base.v[2] + z)))) value = 1;
      }
      if (s->vec[at(s, l, 3, i)] != value) changed++;
      s->vec[at(s, l, 3, i)] = value;
    }
    s->state[20 + l] += changed;
    return;
  }
  case 16:
    WO_PARALLEL for (int i = 0; i < N; i++) {
      int ids[8]; float w[8]; int size = compute_transfers(s, l, i, ids, w);
      for (int k = 0; k < md(s, l, 14); k++) {
        float *e = s->transfer + 2 * ((int64_t)md(s, l, 13) + (int64_t)k * N + i);
        e[0] = k < size ? i2f(ids[k]) : 0; e[1] = k < size ? w[k] : 0;
      }
    }
    return;
  case 2: memset(s->matrix + md(s, l, 4), 0, (size_t)N * 125 * 4); return;
  case 3: { /* pressure_multigrid.glslinc:159-175 */
    int width = md(s, l, 11);
    const int64_t entries = (int64_t)N * width;
    for (int64_t e = 0; e < entries; e++) {
      const int64_t i = s->scatter_reverse ? entries - 1 - e : e;
      int row = (int)(i / width), k = (int)(i % width);
      if (!active_cell(s, l, row)) continue;
      I3 q = coord(row, nl), off = moffset(s, l, k);
      int col = mwrap(s, l, i3(q.v[0] + off.v[0], q.v[1] + off.v[1], q.v[2] + off.v[2]));
      if (col < row || !active_cell(s, l, col)) continue;
      if (l > 0 && !unique_neighbor(s, l, row, off, col)) continue;
      float aij = l == 0 ? (k == 0 ? fine_coefficient(s, 12, row) : -fine_coefficient(s, 5 + k, row)) : mcoefficient(s, l, row, col);
      if (aij == 0) continue;
      int left[8], right[8]; float wl[8], wr[8];
      int nl_ = transfers(s, l, row, left, wl), nr = transfers(s, l, col, right, wr);
      for (int a = 0; a < nl_; a++) for (int b = 0; b < nr; b++) {
        if (row == col && left[a] > right[b]) continue;
        float value = aij * wl[a] * wr[b]; if (row != col && left[a] == right[b]) value *= 2;
        add_entry(s, l + 1, left[a], right[b], value);
      }
    }
    return;
  }
  case 13: {
    int width = md(s, l, 11);
    WO_PARALLEL for (int i = 0; i < N; i++) {
      I3 q = coord(i, nl);
      for (int k = 0; k < width; k++) {
        I3 off = moffset(s, l, k); int j = mwrap(s, l, i3(q.v[0] + off.v[0], q.v[1] + off.v[1], q.v[2] + off.v[2]));
        float a = unique_neighbor(s, l, i, off, j) ? mcoefficient(s, l, i, j) : 0;
        float *e = s->rows + 2 * ((int64_t)md(s, l, 12) + i + (int64_t)k * N); e[0] = a; e[1] = i2f(j > 0 ? j : 0);
      }
    }
    return;
  }
  case 4: /* pressure_multigrid.glslinc:183-189 */
    WO_PARALLEL for (int i = 0; i < N; i++) {
      float sum = 0; I3 q = coord(i, nl);
      if (l == 0) {
        sum = fine_coefficient(s, 12, i);
        for (int axis = 0; axis < 3; axis++) for (int side = 0; side < 2; side++) {
          I3 nb = q; nb.v[axis] += 2 * side - 1; if (active_cell(s, 0, mwrap(s, 0, nb))) sum += fine_coefficient(s, 6 + 2 * axis + side, i);
        }
      } else for (int k = 0; k < md(s, l, 11); k++) {
        I3 off = moffset(s, l, k); int j = mwrap(s, l, i3(q.v[0] + off.v[0], q.v[1] + off.v[1], q.v[2] + off.v[2]));
        if (unique_neighbor(s, l, i, off, j)) sum += fabsf(mcoefficient(s, l, i, j));
      }
      if (!(sum > 0) || isnan(sum) || isinf(sum)) { s->c[7] = 32; continue; }
      s->vec[at(s, l, 4, i)] = 1 / sum;
    }
    return;
  case 5: WO_PARALLEL for (int i = 0; i < N; i++) { s->vec[at(s, l, 0, i)] = s->work[i]; s->vec[at(s, l, 1, i)] = 0; s->vec[at(s, l, 2, i)] = 0; } return;
  case 15: { /* pressure_multigrid.glslinc:192-196 */
    int source = field % 4, k = field / 4;
    WO_PARALLEL for (int i = 0; i < N; i++) {
      float r = s->vec[at(s, l, 0, i)] - apply(s, l, source, i);
      float d = (k == 1 ? 0 : (float)(2 * k - 3) / (float)(2 * k + 1) * s->vec[at(s, l, 7, i)]) + (float)(8 * k - 4) / (float)(2 * k + 1) * s->vec[at(s, l, 4, i)] * r;
      s->vec[at(s, l, 7, i)] = d; s->vec[at(s, l, 3 - source, i)] = s->vec[at(s, l, source, i)] + d;
    }
    return;
  }
  case 14: WO_PARALLEL for (int i = 0; i < N; i++) s->vec[at(s, l, 2, i)] = s->vec[at(s, l, 4, i)] * s->vec[at(s, l, 0, i)]; return;
  case 6: { int source = field, dest = 3 - source; WO_PARALLEL for (int i = 0; i < N; i++) s->vec[at(s, l, dest, i)] = s->vec[at(s, l, source, i)] + s->vec[at(s, l, 4, i)] * (s->vec[at(s, l, 0, i)] - apply(s, l, source, i)); return; }
  case 7: WO_PARALLEL for (int i = 0; i < N; i++) s->vec[at(s, l, 6, i)] = s->vec[at(s, l, 0, i)] - apply(s, l, field, i); return;
  case 8: { /* pressure_multigrid.glslinc:200-224 */
    int parent = l - 1; I3 f = i3(md(s, parent, 5), md(s, parent, 6), md(s, parent, 7)), nf = mdims(s, parent);
    WO_PARALLEL for (int i = 0; i < N; i++) {
      float sum = 0; I3 qc = coord(i, nl);
      if (md(s, parent, 15) != 0) {
        if (active_cell(s, l, i)) for (int z = 0; z < f.v[2]; z++) for (int y = 0; y < f.v[1]; y++) for (int x = 0; x < f.v[0]; x++) {
          int j = mwrap(s, parent, i3(qc.v[0] * f.v[0] + x, qc.v[1] * f.v[1] + y, qc.v[2] * f.v[2] + z));
          if (active_cell(s, parent, j)) sum += s->vec[at(s, parent, 6, j)];
        }
        s->vec[at(s, l, 0, i)] = sum; s->vec[at(s, l, 1, i)] = 0; s->vec[at(s, l, 2, i)] = 0; continue;
      }
      I3 lo, ext = i3(1, 1, 1);
      for (int k = 0; k < 3; k++) lo.v[k] = qc.v[k] * f.v[k];
      for (int axis = 0; axis < 3; axis++) if (f.v[axis] > 1) {
        lo.v[axis]--; ext.v[axis] = 4;
        if ((axis == 2 || (axis == 0 && perx(s))) && nf.v[axis] <= 4) { lo.v[axis] = 0; ext.v[axis] = nf.v[axis]; }
      }
      if (active_cell(s, l, i)) for (int z = 0; z < ext.v[2]; z++) for (int y = 0; y < ext.v[1]; y++) for (int x = 0; x < ext.v[0]; x++) {
        int j = mwrap(s, parent, i3(lo.v[0] + x, lo.v[1] + y, lo.v[2] + z)); if (!active_cell(s, parent, j)) continue;
        I3 q = coord(j, nf); float weight = 1;
        for (int axis = 0; axis < 3; axis++) {
          float pos = ((float)q.v[axis] + .5f) / (float)f.v[axis] - .5f; int base = (int)floorf(pos); float t = fractf(pos);
          float w = 0;
          for (int bit = 0; bit < f.v[axis]; bit++) {
            int nb = base + bit;
            if (axis == 2 || (axis == 0 && perx(s))) nb = pmod(nb, nl.v[axis]);
            else if (axis == 0) nb = clampi(nb, 0, nl.v[axis] - 1); else nb = nb > 0 ? nb : 0;
            if (nb == qc.v[axis]) w += f.v[axis] == 1 ? 1 : (bit == 1 ? t : 1 - t);
          }
          weight *= w;
        }
        sum += weight * s->vec[at(s, parent, 6, j)];
      }
      s->vec[at(s, l, 0, i)] = sum; s->vec[at(s, l, 1, i)] = 0; s->vec[at(s, l, 2, i)] = 0;
    }
    return;
  }
  case 9: {
    int ff = field % 4, cf = field / 4;
    WO_PARALLEL for (int i = 0; i < N; i++) {
      float add = 0;
      if (active_cell(s, l, i)) { int ids[8]; float ws[8]; int size = transfers(s, l, i, ids, ws); for (int k = 0; k < size; k++) add += ws[k] * s->vec[at(s, l + 1, cf, ids[k])]; }
      s->vec[at(s, l, ff, i)] += add;
    }
    return;
  }
  case 10: { /* pressure_multigrid.glslinc:230-234 */
    int64_t G = ((int64_t)N + 127) / 128;
    WO_PARALLEL for (int64_t g = 0; g < G; g++) {
      F4 lane[128];
      for (int lid = 0; lid < 128; lid++) {
        int64_t i = g * 128 + lid; F4 value = f4(0, 0, 0, 0);
        if (i < N) {
          float r = s->work[i], z = s->vec[at(s, 0, field, i)], b = s->work[5 * (int64_t)N + i];
          s->work[N + i] = z; if (phase == 0) s->work[2 * (int64_t)N + i] = z;
          value = f4(r * z, r * r, b * b, fabsf(r));
        }
        lane[lid] = value;
      }
      s->partial[g] = tree_reduce(lane, RED_PRESSURE);
    }
    return;
  }
  }
}

/* pressure_hierarchy_state.glsl:60-68 */
static uint32_t signature(const WO *s, int64_t i) {
  if (s->geo[2 * i + 1] < .5f) return 0;
  I3 n = dims(s); I3 q = coord(i, n); uint32_t result = 0; int64_t N = s->N;
  for (int axis = 0; axis < 3; axis++) for (int side = 0; side < 2; side++) {
    I3 nb = q; nb.v[axis] += 2 * side - 1; nb.v[2] = pmod(nb.v[2], n.v[2]); if (perx(s)) nb.v[0] = pmod(nb.v[0], n.v[0]);
    int inside = nb.v[0] >= 0 && nb.v[0] < n.v[0] && nb.v[1] >= 0 && nb.v[1] < n.v[1];
    int w = inside && s->geo[2 * index3(nb, n) + 1] > .5f;
    uint32_t kind = s->work[(int64_t)(6 + 2 * axis + side) * N + i] > 0 ? (w ? 1u : 2u) : 0u;
    result |= kind << (2 * (2 * axis + side));
  }
  return result;
}
/* pressure_hierarchy_state.glsl:69-109 */
void wo_audit_op(WO *s, int op, int64_t groups) {
  (void)groups;
  if (s->c[5] < .5f || s->c[7] > .5f) return;
  uint32_t *st = s->state; int64_t n = s->N; float *cached = s->snapshot, *prev = s->previous;
  uint32_t capacity = (uint32_t)s->history_capacity;
  if (op == 0) { for (int k = 1; k <= 7; k++) st[k] = 0; st[14] = st[0]; st[15] = 0; st[16] = 0; for (int k = 20; k < 32; k++) st[k] = 0; return; }
  if (op == 1) {
    for (int64_t i = 0; i < n; i++) {
      float kind = s->geo[2 * i + 1];
      st[1] += kind != cached[7 * n + i]; st[4] += kind != prev[7 * n + i]; st[16] += kind > .5f;
      uint32_t sig = signature(s, i), old = (uint32_t)cached[8 * n + i], pv = (uint32_t)prev[8 * n + i];
      for (int face = 0; face < 6; face++) { st[2] += ((sig >> (2 * face)) & 3) != ((old >> (2 * face)) & 3); st[5] += ((sig >> (2 * face)) & 3) != ((pv >> (2 * face)) & 3); }
      float ratio = 0; int bad = 0;
      for (int k = 0; k < 7; k++) {
        float value = s->work[(6 + k) * n + i], was = cached[k * n + i];
        st[3] += f2u(value) != f2u(was); st[6] += f2u(value) != f2u(prev[k * n + i]);
        if (was > 0) ratio = fmaxf_(ratio, fabsf(value - was) / was);
        if (isnan(value) || isinf(value) || value < 0) bad = 1;
      }
      if (f2u(ratio) > st[7]) st[7] = f2u(ratio);
      st[15] += bad;
    }
    return;
  }
  if (op == 2) {
    if (st[15] > 0) { s->c[7] = 34; return; }
    int mode = s->mode;
    int incompatible = st[0] == 0 || st[1] > 0 || st[2] > 0;
    int structural = mode == 0 || incompatible;
    int numerical = mode == 0 || incompatible || (st[3] > 0 && (mode == 1 || st[10] + 1 >= (uint32_t)s->lag));
    st[8] = numerical; st[9] = structural; st[10] = numerical || st[3] == 0 ? 0 : st[10] + 1; st[11] += st[8]; st[12] += st[9]; st[13]++; st[0] = 1;
    if (st[13] <= capacity) {
      int64_t row = 64 + 32 * ((int64_t)st[13] - 1); st[row] = st[13]; st[row + 1] = f2u(s->c[1]);
      for (int k = 1; k <= 7; k++) st[row + 1 + k] = st[k];
      st[row + 9] = st[8]; st[row + 10] = st[9]; st[row + 11] = st[10]; st[row + 12] = st[16]; st[row + 13] = st[11];
    }
    return;
  }
  if (op == 3) {
    WO_PARALLEL for (int64_t i = 0; i < n; i++) {
      for (int k = 0; k < 7; k++) { float value = s->work[(6 + k) * n + i]; prev[k * n + i] = value; if (st[8] != 0) cached[k * n + i] = value; }
      prev[7 * n + i] = s->geo[2 * i + 1]; prev[8 * n + i] = (float)signature(s, i);
      if (st[8] != 0) { cached[7 * n + i] = prev[7 * n + i]; cached[8 * n + i] = prev[8 * n + i]; }
    }
    return;
  }
  if (op == 4) {
    if (st[13] <= capacity) { int64_t row = 64 + 32 * ((int64_t)st[13] - 1); for (int l = 0; l < 12; l++) st[row + 16 + l] = st[20 + l]; }
    return;
  }
}

/* ---- level planning: water/pressure_multigrid.gd:33-134, 135-154, 181-198 ---- */
typedef struct { int v[3]; } Off;
static int floor_div_basis(int q, int f) { return (int)floorf(((float)q + .5f) / (float)f - .5f); }
static int basis(I3 q, I3 f, int aggregate, I3 *out) {
  if (aggregate) { for (int k = 0; k < 3; k++) out[0].v[k] = (int)floorf((float)q.v[k] / (float)f.v[k]); return 1; }
  I3 base; for (int k = 0; k < 3; k++) base.v[k] = floor_div_basis(q.v[k], f.v[k]);
  int m = 0;
  for (int z = 0; z < f.v[2]; z++) for (int y = 0; y < f.v[1]; y++) for (int x = 0; x < f.v[0]; x++) out[m++] = i3(base.v[0] + x, base.v[1] + y, base.v[2] + z);
  return m;
}
static int coarse_stencil(const I3 *st, int ns, I3 f, int aggregate, I3 *out) {
  int m = 0;
  for (int z = 0; z < f.v[2]; z++) for (int y = 0; y < f.v[1]; y++) for (int x = 0; x < f.v[0]; x++) {
    I3 q = i3(x, y, z), left[8]; int nl = basis(q, f, aggregate, left);
    for (int o = 0; o < ns; o++) for (int a = 0; a < nl; a++) {
      I3 right[8]; int nr = basis(i3(q.v[0] + st[o].v[0], q.v[1] + st[o].v[1], q.v[2] + st[o].v[2]), f, aggregate, right);
      for (int b = 0; b < nr; b++) {
        I3 d = i3(right[b].v[0] - left[a].v[0], right[b].v[1] - left[a].v[1], right[b].v[2] - left[a].v[2]);
        int found = 0; for (int e = 0; e < m; e++) if (out[e].v[0] == d.v[0] && out[e].v[1] == d.v[1] && out[e].v[2] == d.v[2]) { found = 1; break; }
        if (!found) out[m++] = d;
      }
    }
  }
  return m;
}
static void add_cmd(int (**list)[4], int *n, int op, int l, int field, int kind) {
  *list = realloc(*list, sizeof(int[4]) * (*n + 1)); (*list)[*n][0] = op; (*list)[*n][1] = l; (*list)[*n][2] = field; (*list)[*n][3] = kind; (*n)++;
}
static int smooth_cmds(WO *s, int l, int field) {
  for (int sweep = 0; sweep < s->smoothing; sweep++) { add_cmd(&s->cycle, &s->ncycle, 15, l, field + 4 * (sweep + 1), 0); field = 3 - field; }
  return field;
}
static int cycle_cmds(WO *s, int l) {
  if (l == s->L - 1) {
    if (mcount(s, l) <= 64) { add_cmd(&s->cycle, &s->ncycle, 12, l, 0, -1); return 1; }
    int tf = 1;
    for (int sweep = 0; sweep < 32; sweep++) { add_cmd(&s->cycle, &s->ncycle, sweep == 0 ? 14 : 6, l, tf, 0); tf = 3 - tf; }
    return tf;
  }
  int field = smooth_cmds(s, l, 1);
  add_cmd(&s->cycle, &s->ncycle, 7, l, field, 0); add_cmd(&s->cycle, &s->ncycle, 8, l + 1, 0, 0);
  int cf = cycle_cmds(s, l + 1);
  add_cmd(&s->cycle, &s->ncycle, 9, l, field + 4 * cf, 0);
  return smooth_cmds(s, l, field);
}

int wo_mg_init(WO *s) {
  I3 n = dims(s); F3 h = f3(PC_H(s, 0), PC_H(s, 1), PC_H(s, 2));
  static I3 stencils[MAXL][125]; int ns[MAXL];
  I3 st[125]; int nst = 7;
  st[0] = i3(0, 0, 0); st[1] = i3(-1, 0, 0); st[2] = i3(1, 0, 0); st[3] = i3(0, -1, 0); st[4] = i3(0, 1, 0); st[5] = i3(0, 0, -1); st[6] = i3(0, 0, 1);
  int meta[16 * MAXL]; memset(meta, 0, sizeof meta);
  int64_t node_count = 0, matrix_count = 0; int L = 0;
  for (;;) {
    int count = n.v[0] * n.v[1] * n.v[2]; I3 f = i3(1, 1, 1);
    if (count > 64) {
      float smallest = INFINITY;
      for (int ax = 0; ax < 3; ax++) if (n.v[ax] > 2) smallest = fminf_(smallest, h.v[ax]);
      for (int ax = 0; ax < 3; ax++) if (n.v[ax] > 2 && h.v[ax] < 1.8f * smallest) f.v[ax] = 2;
      if (perx(s) && n.v[0] % 2) f.v[0] = 1;
      if (n.v[2] % 2) f.v[2] = 1;
    }
    int aggregate = L >= s->aggregate_from;
    if (L >= MAXL) return -1;
    memcpy(stencils[L], st, sizeof(I3) * nst); ns[L] = nst;
    int *m = meta + 16 * L;
    m[0] = n.v[0]; m[1] = n.v[1]; m[2] = n.v[2]; m[3] = (int)node_count; m[4] = (int)matrix_count; m[5] = f.v[0]; m[6] = f.v[1]; m[7] = f.v[2]; m[8] = count; m[15] = aggregate;
    L++;
    node_count += count;
    if (L > 1) matrix_count += (int64_t)count * 125;
    if (count <= 64 || (f.v[0] == 1 && f.v[1] == 1 && f.v[2] == 1)) break;
    I3 next[125]; int nn = coarse_stencil(st, nst, f, aggregate, next);
    memcpy(st, next, sizeof(I3) * nn); nst = nn;
    for (int k = 0; k < 3; k++) { n.v[k] = (n.v[k] + f.v[k] - 1) / f.v[k]; h.v[k] *= (float)f.v[k]; }
  }
  s->L = L;
  int total_off = 0; for (int l = 0; l < L; l++) total_off += ns[l];
  s->nmeta = 16 * L + 3 * total_off;
  s->meta = calloc((size_t)s->nmeta, sizeof(int));
  memcpy(s->meta, meta, sizeof(int) * 16 * L);
  int pos = 16 * L; int64_t row_count = 0, transfer_count = 0;
  for (int l = 0; l < L; l++) {
    int *m = s->meta + 16 * l;
    m[10] = pos; m[11] = ns[l]; m[12] = (int)row_count; m[13] = (int)transfer_count;
    m[14] = m[15] ? 1 : m[5] * m[6] * m[7];
    if (l + 1 < L) transfer_count += (int64_t)m[8] * m[14];
    if (l > 0) row_count += (int64_t)m[8] * ns[l];
    for (int k = 0; k < ns[l]; k++) { s->meta[pos++] = stencils[l][k].v[0]; s->meta[pos++] = stencils[l][k].v[1]; s->meta[pos++] = stencils[l][k].v[2]; }
  }
  s->meta[16 * (L - 1) + 9] = (int)matrix_count;
  matrix_count += (int64_t)s->meta[16 * (L - 1) + 8] * s->meta[16 * (L - 1) + 8];
  s->matrix_count = matrix_count; s->node_count = node_count; s->row_count = row_count; s->transfer_count = transfer_count;
  s->matrix = calloc((size_t)matrix_count, 4);
  s->vec = calloc((size_t)node_count * 8, 4);
  s->rows = calloc((size_t)(row_count > 0 ? row_count : 1) * 2, 4);
  s->transfer = calloc((size_t)(transfer_count > 0 ? transfer_count : 1) * 2, 4);
  s->snapshot = calloc((size_t)s->N * 9, 4); s->previous = calloc((size_t)s->N * 9, 4);
  s->state_count = 64 + 32 * (int64_t)s->history_capacity;
  s->state = calloc((size_t)s->state_count, 4);
  /* build_commands, water/pressure_multigrid.gd:121-128 */
  add_cmd(&s->build, &s->nbuild, 0, 0, 0, -1);
  for (int l = 0; l < L; l++) add_cmd(&s->build, &s->nbuild, 1, l, 0, 3);
  add_cmd(&s->build, &s->nbuild, 0, 0, 0, -1);
  for (int l = 0; l < L - 1; l++) add_cmd(&s->build, &s->nbuild, 16, l, 0, 4);
  add_cmd(&s->build, &s->nbuild, 4, 0, 0, 0);
  for (int l = 1; l < L; l++) { add_cmd(&s->build, &s->nbuild, 2, l, 0, 1); add_cmd(&s->build, &s->nbuild, 3, l - 1, 0, 2); add_cmd(&s->build, &s->nbuild, 4, l, 0, 0); }
  for (int l = 1; l < L; l++) add_cmd(&s->build, &s->nbuild, 13, l, 0, 0);
  if (s->meta[16 * (L - 1) + 8] <= 64) add_cmd(&s->build, &s->nbuild, 11, L - 1, 0, 5);
  /* cycle_commands, water/pressure_multigrid.gd:129-131, 135-154 */
  add_cmd(&s->cycle, &s->ncycle, 0, 0, 0, -1); add_cmd(&s->cycle, &s->ncycle, 5, 0, 0, 0);
  int final_field = cycle_cmds(s, 0);
  add_cmd(&s->cycle, &s->ncycle, 10, 0, final_field, 0);
  return 0;
}
void wo_mg_free(WO *s) {
  free(s->meta); free(s->matrix); free(s->vec); free(s->rows); free(s->transfer); free(s->snapshot); free(s->previous); free(s->state); free(s->cycle); free(s->build);
}

/* water/pressure_multigrid.gd:155-176 */
void wo_mg_record(WO *s, int build, int initial) {
  if (build) for (int op = 0; op < 4; op++) wo_audit_op(s, op, 0);
  int phase = build ? 2 : (initial ? 0 : 1);
  int (*list)[4] = build ? s->build : s->cycle; int n = build ? s->nbuild : s->ncycle;
  for (int k = 0; k < n; k++) { if (list[k][0] == 0) continue; wo_mg_op(s, list[k][0], list[k][1], phase, list[k][2], 0); }
  if (build) wo_audit_op(s, 4, 0);
}

int wo_levels(WO *s) { return s->L; }
int wo_level_meta(WO *s, int level, int *dst16) { if (level < 0 || level >= s->L) return -1; memcpy(dst16, s->meta + 16 * level, 64); return 0; }

/* The V-cycle as an operator, outside the control gating (tests only). */
void wo_precondition(WO *s, const float *r, float *z) {
  float saved[NCTRL]; memcpy(saved, s->c, sizeof saved);
  float *w0 = malloc(sizeof(float) * s->N * 6); memcpy(w0, s->work, sizeof(float) * s->N * 6);
  memcpy(s->work, r, sizeof(float) * s->N);
  s->c[4] = 1; s->c[5] = 1; s->c[6] = 0; s->c[7] = 0; s->c[57] = 0;
  wo_mg_record(s, 0, 0);
  memcpy(z, s->work + s->N, sizeof(float) * s->N);
  memcpy(s->work, w0, sizeof(float) * s->N * 6); free(w0);
  memcpy(s->c, saved, sizeof saved);
}
