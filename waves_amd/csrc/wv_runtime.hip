// Host runtime and C-ABI of the gfx950 solver (libwaves_hip.so).
//
// Replaces the reference's render-thread runtime: buffer allocation
// (water/pressure_gpu.gd:82-127, water/coupled_gpu.gd:23-33), multigrid
// planning (water/pressure_multigrid.gd:33-134), the per-dispatch driver
// (water/pressure_gpu.gd:69-81) and the chunk scheduler with one asynchronous
// control readback per submission (water/work_scheduler.gd:41-87).
// A batch of the coupled step program is captured once into a hipGraph and
// replayed; every op is gated on the device-resident control vector, so the
// host never decides acceptance, retries or PCG convergence.
#include <hip/hip_runtime.h>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include "wv_kernels.h"
#include "../../include/waves.h"
#include "../../include/waves_net.h"
#include "wv_netdev.h"

namespace {

struct Offset { int x, y, z; };

struct Plan {
  int L = 0;
  MGLevel lv[WV_MAXL];
  std::vector<int> meta;            // 16 ints per level, the reference layout
  std::vector<int> offsets;         // 3 ints per stencil entry
  std::vector<signed char> slots;   // 125 per level
  std::vector<int> deltas;          // 1 int per stencil entry
  int64_t node_count = 0, row_count = 0, ap_floats = 0;
};

// water/pressure_multigrid.gd:181-188
static int basis(Offset q, const int f[3], bool aggregate, Offset* out) {
  if (aggregate) { out[0] = {(int)std::floor((float)q.x / f[0]), (int)std::floor((float)q.y / f[1]), (int)std::floor((float)q.z / f[2])}; return 1; }
  int b[3], qa[3] = {q.x, q.y, q.z};
  for (int k = 0; k < 3; k++) b[k] = (int)std::floor(((double)qa[k] + .5) / f[k] - .5);
  int m = 0;
  for (int z = 0; z < f[2]; z++) for (int y = 0; y < f[1]; y++) for (int x = 0; x < f[0]; x++) out[m++] = {b[0] + x, b[1] + y, b[2] + z};
  return m;
}
// water/pressure_multigrid.gd:189-198 (insertion order kept: it fixes the row layout)
static std::vector<Offset> coarse_stencil(const std::vector<Offset>& st, const int f[3], bool aggregate) {
  std::vector<Offset> out;
  for (int z = 0; z < f[2]; z++) for (int y = 0; y < f[1]; y++) for (int x = 0; x < f[0]; x++) {
    Offset q{x, y, z}, left[8]; int nl = basis(q, f, aggregate, left);
    for (const Offset& o : st) for (int a = 0; a < nl; a++) {
      Offset right[8]; int nr = basis({q.x + o.x, q.y + o.y, q.z + o.z}, f, aggregate, right);
      for (int b = 0; b < nr; b++) {
        Offset d{right[b].x - left[a].x, right[b].y - left[a].y, right[b].z - left[a].z};
        bool found = false; for (const Offset& e : out) if (e.x == d.x && e.y == d.y && e.z == d.z) { found = true; break; }
        if (!found) out.push_back(d);
      }
    }
  }
  return out;
}
// water/pressure_multigrid.gd:40-69
static bool plan_levels(const float* pc, int aggregate_from, Plan& p) {
  int n[3] = {(int)pc[0], (int)pc[1], (int)pc[2]}; float h[3] = {pc[4], pc[5], pc[6]};
  bool perx = pc[12] > .5f;
  std::vector<Offset> st = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  std::vector<std::vector<Offset>> stencils;
  int64_t node_count = 0, matrix_count = 0;
  for (;;) {
    int count = n[0] * n[1] * n[2]; int f[3] = {1, 1, 1};
    if (count > 64) {
      float smallest = INFINITY;
      for (int a = 0; a < 3; a++) if (n[a] > 2) smallest = std::fmin(smallest, h[a]);
      for (int a = 0; a < 3; a++) if (n[a] > 2 && h[a] < 1.8f * smallest) f[a] = 2;
      if (perx && n[0] % 2) f[0] = 1;
      if (n[2] % 2) f[2] = 1;
    }
    if (p.L >= WV_MAXL) return false;
    bool aggregate = p.L >= aggregate_from;
    MGLevel& L = p.lv[p.L];
    for (int a = 0; a < 3; a++) { L.n[a] = n[a]; L.f[a] = f[a]; }
    L.count = count; L.aggregate = aggregate; L.width = (int)st.size();
    L.vec_off = 8 * node_count;
    int m[16] = {n[0], n[1], n[2], (int)node_count, (int)matrix_count, f[0], f[1], f[2], count, 0, 0, 0, 0, 0, 0, aggregate ? 1 : 0};
    p.meta.insert(p.meta.end(), m, m + 16);
    stencils.push_back(st);
    p.L++;
    node_count += count;
    if (p.L > 1) matrix_count += (int64_t)count * 125;
    if (count <= 64 || (f[0] == 1 && f[1] == 1 && f[2] == 1)) break;
    st = coarse_stencil(st, f, aggregate);
    for (int a = 0; a < 3; a++) { n[a] = (n[a] + f[a] - 1) / f[a]; h[a] *= (float)f[a]; }
  }
  int64_t row_count = 0, transfer_count = 0; int pos = 16 * p.L;
  p.slots.assign(125 * p.L, (signed char)-1);
  for (int l = 0; l < p.L; l++) {
    MGLevel& L = p.lv[l]; int* m = &p.meta[16 * l];
    m[10] = pos; m[11] = L.width; m[12] = (int)row_count; m[13] = (int)transfer_count;
    m[14] = L.aggregate ? 1 : L.f[0] * L.f[1] * L.f[2];
    L.off_start = (int)(p.offsets.size() / 3);
    L.row_off = l > 0 ? row_count : 0;
    if (l + 1 < p.L) transfer_count += (int64_t)L.count * m[14];
    if (l > 0) row_count += (int64_t)L.count * L.width;
    L.margin[0] = L.margin[1] = L.margin[2] = 0;
    for (int k = 0; k < L.width; k++) {
      const Offset& o = stencils[l][k];
      L.margin[0] = std::max(L.margin[0], std::abs(o.x)); L.margin[1] = std::max(L.margin[1], std::abs(o.y)); L.margin[2] = std::max(L.margin[2], std::abs(o.z));
      p.deltas.push_back(o.x + L.n[0] * (o.y + L.n[1] * o.z));
      p.offsets.push_back(o.x); p.offsets.push_back(o.y); p.offsets.push_back(o.z);
      if (std::abs(o.x) <= 2 && std::abs(o.y) <= 2 && std::abs(o.z) <= 2) p.slots[125 * l + (o.x + 2 + 5 * (o.y + 2 + 5 * (o.z + 2)))] = (signed char)k;
      pos += 3;
    }
    L.wrap1 = L.margin[2] <= L.n[2] && (!perx || L.margin[0] <= L.n[0]);
  }
  // Galerkin stage-1 boxes: coarse span of the parents of a fine row's stencil
  p.ap_floats = 0;
  for (int l = 0; l + 1 < p.L; l++) {
    MGLevel& F = p.lv[l];
    int64_t e = 1;
    for (int a = 0; a < 3; a++) {
      int mx = 0, m = F.margin[a], f = F.f[a];
      for (int q = 0; q < 2 * f + 2; q++) {
        auto pb = [&](int j) -> int {
          if (F.aggregate) return (int)std::floor((double)j / f);
          if (f == 1) return j;
          return (int)std::floor(((float)j + .5f) / (float)f - .5f);
        };
        int span = pb(q + m) - pb(q - m) + (F.aggregate || f == 1 ? 1 : 2);
        mx = std::max(mx, span);
      }
      F.box[a] = mx; e *= mx;
    }
    if (e > 64) return false;  // stage-1 LDS holds 64 box entries per fine row
    p.ap_floats = std::max(p.ap_floats, ((e + 3) / 4 * 4) * F.count);  // rows padded to 4 floats
  }
  // lanes per row: enough threads that a sweep's load chains overlap on the
  // small levels (about 2^18 threads); a level with more rows than that keeps
  // one row per thread, whose unit-stride loop issues its loads back to back
  // (measured faster than 4 lanes per row on the 297,600-row level 1)
  const char* lt = getenv("WAVES_MG_LANE_TARGET");
  const int64_t lane_target = lt ? std::atoll(lt) : (1 << 18);
  for (int l = 0; l < p.L; l++) {
    MGLevel& L = p.lv[l]; int g = 1;
    while (g < 32 && (int64_t)L.count * g < lane_target && g * 2 <= L.width) g *= 2;
    L.lanes = l == 0 ? 1 : g;
    // the restriction gathers a 4x4x4 (trilinear) or 2x2x2 box per node with
    // no unrolled stencil: it keeps about 2^20 lanes on every level
    int box = 1;
    if (l > 0) for (int a = 0; a < 3; a++) box *= p.lv[l - 1].aggregate ? p.lv[l - 1].f[a] : (p.lv[l - 1].f[a] > 1 ? 4 : 1);
    int r = 1;
    while (r < 32 && 2 * r <= box && (int64_t)L.count * r < (1 << 20)) r *= 2;
    L.rlanes = r;
    // Galerkin stage 1 (A P rows of this level's fine rows): about the same
    // lane target, at most 16 lanes and the row width.  (Four lanes per row on
    // the 297,600-row level 1 take 108 us against 154 us, but the lane split
    // changes the order of the row sums and so the trajectory; kept at one.)
    int ga = 1;
    while (ga < 16 && (int64_t)L.count * ga < lane_target && 2 * ga <= L.width) ga *= 2;
    L.ga = ga;
  }
  p.meta[16 * (p.L - 1) + 9] = (int)matrix_count;
  p.node_count = node_count; p.row_count = row_count;
  p.ap_floats = std::max(p.ap_floats, (int64_t)p.lv[0].count);  // also the copy kernel's destination
  return true;
}

}  // namespace

struct wv_solver {
  wv_settings settings;
  Ctx ctx;
  MGDev mg;
  Plan plan;
  hipStream_t stream = nullptr;
  std::vector<void*> allocations;
  float* quota_dev = nullptr;
  float* quota_host = nullptr;   // pinned
  float* ctrl_host = nullptr;    // pinned, one 96-float readback slot per submission in flight
  float* ctl_timing = nullptr;   // device control vector used only by wv_time_kernel
  float* ctl_saved = nullptr;    // the caller's control vector around a standalone pressure solve
  wv_net* net = nullptr;         // learned warm start (not owned)
  // submissions in flight (water/work_scheduler.gd:48-54: at most three), oldest first
  struct Submission { hipEvent_t start = nullptr, end = nullptr; int batches = 0; double record_us = 0; };
  Submission sub[WV_MAX_INFLIGHT];
  int sub_head = 0, sub_count = 0;
  std::atomic<double> quota{0.0};
  float last_uploaded_quota = NAN;
  float last_control[WV_NCTRL] = {0};
  wv_schedule_stats sched{};     // counters of the scheduler (telemetry of water/work_scheduler.gd:18-40)
  double ema_batch_ms = 0;       // EMA of one batch's GPU time (the reference's per-chunk cost EMA)
  hipGraphExec_t batch_exec = nullptr;
  hipGraph_t batch_graph = nullptr;
  // the iterations the last standalone solve of each form needed (the next
  // solve's first chunk)
  int solve_iterations[3] = {0, 0, 0};
  int smoothing = 3, iterations = 8;
  int precision = 0;      // batch program: 0 the reference's FP32 PCG, 1 BF16 V-cycle + FP64 dots
  bool use_graph = true;  // WAVES_HIP_GRAPH=0: launch each batch's kernels directly (tracing tools)
  int64_t history_capacity = 0;
  std::string error;
  std::mutex lock;
};

static int fail(wv_solver* s, int code, const std::string& msg) {
  if (s) s->error = msg;
  return code;
}
#define HIPCHECK(s, expr)                                                           \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) return fail((s), WV_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static int dalloc(wv_solver* s, T** ptr, size_t count) {
  void* p = nullptr;
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = 16;
  if (hipMalloc(&p, bytes) != hipSuccess) return fail(s, WV_ENOMEM, "hipMalloc failed for " + std::to_string(bytes) + " bytes");
  if (hipMemset(p, 0, bytes) != hipSuccess) return fail(s, WV_EHIP, "hipMemset failed");
  s->allocations.push_back(p);
  *ptr = (T*)p;
  return 0;
}

static void release(wv_solver* s) {
  if (s->batch_exec) hipGraphExecDestroy(s->batch_exec);
  if (s->batch_graph) hipGraphDestroy(s->batch_graph);
  if (s->stream) hipStreamSynchronize(s->stream);
  for (auto it = s->allocations.rbegin(); it != s->allocations.rend(); ++it) hipFree(*it);
  s->allocations.clear();
  if (s->quota_host) hipHostFree(s->quota_host);
  if (s->ctrl_host) hipHostFree(s->ctrl_host);
  for (auto& b : s->sub) {
    if (b.start) hipEventDestroy(b.start);
    if (b.end) hipEventDestroy(b.end);
  }
  if (s->stream) hipStreamDestroy(s->stream);
}

extern "C" int wv_create(const wv_settings* settings, wv_solver** out) {
  if (!settings || !out) return WV_EINVAL;
  *out = nullptr;
  const float* pc = settings->pc;
  if (pc[0] < 1 || pc[1] < 1 || pc[2] < 1 || pc[4] <= 0 || pc[5] <= 0 || pc[6] <= 0) return WV_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return WV_ENODEV;
  wv_solver* s = new wv_solver();
  s->settings = *settings;
  Ctx& c = s->ctx;
  memset(&c, 0, sizeof c);
  c.nx = (int)pc[0]; c.ny = (int)pc[1]; c.nz = (int)pc[2];
  c.N = (int64_t)c.nx * c.ny * c.nz;
  c.NU = (int64_t)(c.nx + 1) * c.ny * c.nz; c.NV = (int64_t)c.nx * (c.ny + 1) * c.nz; c.NW = (int64_t)c.nx * c.ny * (c.nz + 1);
  c.NF = c.NU + c.NV + c.NW; c.stride = 2 * c.N + c.NF;
  for (int a = 0; a < 3; a++) { c.h[a] = pc[4 + a]; c.o[a] = pc[8 + a]; }
  c.dt_max = pc[7]; c.eta = pc[11]; c.perx = pc[12]; c.slope = pc[13]; c.corr = pc[14];
  c.shape = pc[16]; c.gmult = pc[17]; c.max_face_cfl = pc[18]; c.hist_cap = pc[19];
  for (int k = 0; k < 4; k++) { c.wave[k] = pc[20 + k]; c.forc[k] = pc[28 + k]; }
  c.tol = pc[24]; c.linear_air = pc[25]; c.stokes2 = pc[26]; c.flags = (int)pc[27];
  const float* o = settings->opts;
  c.right_abs = o[0]; c.pert_amp = o[6]; c.seed = (uint32_t)o[7];
  s->smoothing = (int)o[1] > 0 ? (int)o[1] : 3;
  s->iterations = (int)o[5] > 0 ? (int)o[5] : 8;
  s->history_capacity = (int64_t)pc[19];
  { const char* e = getenv("WAVES_HIP_GRAPH"); if (e && e[0] == '0') s->use_graph = false; }
  if (!plan_levels(pc, (int)o[3], s->plan)) { delete s; return WV_EPLAN; }
  int rc;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) { delete s; return WV_EHIP; }
#define ALLOC(ptr, n) if ((rc = dalloc(s, &(ptr), (size_t)(n))) != 0) { release(s); delete s; return rc; }
  ALLOC(c.bank[0], c.stride); ALLOC(c.bank[1], c.stride);
  ALLOC(c.work, 14 * c.N);
  ALLOC(c.geo, c.N);
  ALLOC(c.c, WV_NCTRL);
  ALLOC(c.plane, c.N);
  ALLOC(c.motion, 4 * c.N + 12 * c.NF);
  ALLOC(c.history, s->history_capacity * 4 + 4);
  ALLOC(c.chist, s->history_capacity * 5 + 5);
  ALLOC(c.probe, 8);
  ALLOC(c.img_planes, c.N);
  ALLOC(c.img_kinds, c.N);
  ALLOC(c.partial, (c.stride + WV_RED - 1) / WV_RED + 1);
  ALLOC(c.partial64, 2 * ((c.N + WV_RED - 1) / WV_RED + 1));
  ALLOC(c.dctl, 16);
  ALLOC(c.pacc, c.N);
  ALLOC(c.aux, 2 * c.N);
  ALLOC(c.vpart, 4 * ((c.N + WV_RED - 1) / WV_RED) + 4);
  ALLOC(c.mq, c.NF + 8 * 256);   // the face count plus eight blocks (block-aggregated sub-queues)
  ALLOC(c.mq_count, 4 + 4 * 8);  // per queue: its length, then eight sub-queue counters
  ALLOC(c.reflect, 3 * (int64_t)(c.nx + 1) * (c.ny + 1));
  ALLOC(c.tstate, 8);
  ALLOC(c.step_times, 4 * (s->history_capacity + 1));
  ALLOC(s->ctl_timing, WV_NCTRL);
  ALLOC(s->ctl_saved, WV_NCTRL);
  ALLOC(s->quota_dev, 4);
  c.quota = s->quota_dev;
  MGDev& m = s->mg;
  memset(&m, 0, sizeof m);
  m.L = s->plan.L; m.perx = c.perx > .5f;
  for (int l = 0; l < m.L; l++) m.lv[l] = s->plan.lv[l];
  { const char* e = getenv("WAVES_MG_TAIL"); m.tail = (e && e[0] == '0') ? -1 : wv_mg_tail_level(m); }
  int* offs; signed char* slots; int* deltas;
  ALLOC(offs, s->plan.offsets.size());
  ALLOC(deltas, s->plan.deltas.size());
  ALLOC(slots, s->plan.slots.size());
  ALLOC(m.vec, 8 * s->plan.node_count);
  ALLOC(m.rows, s->plan.row_count);
  int last = s->plan.lv[m.L - 1].count;
  ALLOC(m.linv, (int64_t)last * last);
  ALLOC(m.ap, s->plan.ap_floats);
  ALLOC(m.rows16, s->plan.row_count);
  ALLOC(m.snap16, 7 * c.N);
  ALLOC(m.snapshot, 9 * c.N);
  ALLOC(m.previous, 9 * c.N);
  ALLOC(m.state, 64 + 32 * s->history_capacity);
  ALLOC(m.audit_part, 64 * 16);
  { const int64_t ts = wv_mg_tail_table_shorts(m); if (ts > 0) ALLOC(m.ttab, ts); }
#undef ALLOC
  m.offsets = offs; m.slots = slots; m.deltas = deltas;
  m.lag = (int)o[2]; m.mode = (int)o[4]; m.capacity = (int)s->history_capacity;
  if (hipMemcpy(offs, s->plan.offsets.data(), s->plan.offsets.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(slots, s->plan.slots.data(), s->plan.slots.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(deltas, s->plan.deltas.data(), s->plan.deltas.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) { release(s); delete s; return WV_EHIP; }
  if (hipHostMalloc((void**)&s->quota_host, 64) != hipSuccess ||
      hipHostMalloc((void**)&s->ctrl_host, WV_MAX_INFLIGHT * WV_NCTRL * sizeof(float)) != hipSuccess) { release(s); delete s; return WV_EHIP; }
  for (auto& b : s->sub)
    if (hipEventCreate(&b.start) != hipSuccess || hipEventCreate(&b.end) != hipSuccess) { release(s); delete s; return WV_EHIP; }
  memset(s->ctrl_host, 0, WV_MAX_INFLIGHT * WV_NCTRL * sizeof(float));
  {  // device wall clock (pressure-section timing): kHz -> ns per tick
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
    c.tick_ns = (float)(1e6 / khz);
  }
  wv_launch_reflect_table(c, s->stream);
  wv_mg_tail_tables(m, m.ttab, s->stream);
  if (hipStreamSynchronize(s->stream) != hipSuccess) { release(s); delete s; return WV_EHIP; }
  *out = s;
  return WV_OK;
}

extern "C" int wv_destroy(wv_solver* s) {
  if (!s) return WV_EINVAL;
  release(s);
  delete s;
  return WV_OK;
}

extern "C" const char* wv_last_error(wv_solver* s) { return s ? s->error.c_str() : "null solver"; }

// Upload the quota (stream-ordered, outside any graph) when it changed.
static int sync_quota(wv_solver* s) {
  float q = (float)s->quota.load();
  if (q == s->last_uploaded_quota) return 0;
  HIPCHECK(s, hipStreamSynchronize(s->stream));  // the pinned slot may still be in flight
  s->quota_host[0] = q;
  HIPCHECK(s, hipMemcpyAsync(s->quota_dev, s->quota_host, sizeof(float), hipMemcpyHostToDevice, s->stream));
  s->last_uploaded_quota = q;
  return 0;
}

// water/pressure_gpu.gd:69-81: one op; the MG hooks are explicit in the batch.
static int dispatch(wv_solver* s, int op) {
  const Ctx& c = s->ctx;
  if (op == WV_OP_SPMV_DOT) { wv_launch_spmv_dot(c, s->stream); return 0; }
  if (op == 3) {
    wv_launch_op(c, 3, s->stream);
    wv_launch_pcg_init_partials(c, s->stream);
    return 0;
  }
  if (op == 8) { wv_launch_pcg_update(c, s->mg, s->stream); return 0; }
  int rc = wv_launch_op(c, op, s->stream);
  if (rc != 0) return fail(s, WV_EINVAL, "unknown op " + std::to_string(op));
  return 0;
}

// water/coupled_gpu.gd:38-56 + water/coastal_gpu.gd:28-33 with the MG hooks of
// water/pressure_gpu.gd:70-74.
static int record_batch(wv_solver* s) {
  auto d = [&](int op) { return dispatch(s, op); };
  d(43); d(1);
  d(20); d(21); d(45); d(46); d(22);
  // ops 25 + 26 of each sweep; the finalisers' in-order ledger sums are added
  // by op 12's finaliser, all sweeps' and the volume's chains side by side
  for (int sweep = 0; sweep < 3; sweep++) { d(23); d(24); d(55); d(56); wv_launch_vof_deferred(s->ctx, sweep, s->stream); }
  d(27); d(28);
  for (int op : {80, 81, 82, 83, 84, 87, 88}) d(op);
  d(17); d(18); d(16);
  d(2); d(47); d(63); d(64);
  wv_launch_press_begin(s->ctx, s->stream);
  wv_mg_build(s->ctx, s->mg, s->stream);
  const bool mixed = s->precision == 1;
  if (mixed) wv_mg_to_bf16(s->ctx, s->mg, s->stream);
  d(3);
  if (mixed) wv_mg_cycle_bf16(s->ctx, s->mg, 0, s->smoothing, true, true, s->stream);  // + op 4 in FP64
  else wv_mg_cycle(s->ctx, s->mg, 0, s->smoothing, true, s->stream);                  // + op 4
  for (int it = 0; it < s->iterations; it++) {
    if (mixed) { d(5); wv_solve_dot_alpha64(s->ctx, s->stream); }   // ops 6 + 7 in FP64
    else wv_launch_spmv_dot(s->ctx, s->stream);                      // ops 5 + 6 + 7
    d(8);
    if (mixed) wv_mg_cycle_bf16(s->ctx, s->mg, 1, s->smoothing, false, true, s->stream);  // + op 9 in FP64
    else wv_mg_cycle(s->ctx, s->mg, 1, s->smoothing, false, s->stream);                  // + op 9
    d(10);
  }
  d(11); d(65);
  wv_launch_press_end(s->ctx, s->stream);
  d(47);
  for (int op : {30, 31, 32, 33, 34, 47, 48}) d(op);
  d(35); d(36);
  d(62); d(66); wv_launch_accept(s->ctx, true, s->stream); d(13);
  wv_launch_step_record(s->ctx, s->stream);
  d(44); d(19);
  d(85); d(89);
  return 0;
}

static int ensure_graph(wv_solver* s) {
  if (s->batch_exec) return 0;
  HIPCHECK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  record_batch(s);
  hipGraph_t g = nullptr;
  HIPCHECK(s, hipStreamEndCapture(s->stream, &g));
  s->batch_graph = g;
  HIPCHECK(s, hipGraphInstantiate(&s->batch_exec, g, nullptr, nullptr, 0));
  return 0;
}

extern "C" int wv_reset(wv_solver* s) {
  if (!s) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  s->quota.store(0.0);
  int rc = sync_quota(s); if (rc) return rc;
  // the reference allocates cleared buffers (water/pressure_gpu.gd:99); a reset
  // here clears the same state in place
  HIPCHECK(s, hipMemsetAsync(s->ctx.work, 0, 14 * s->ctx.N * sizeof(float), s->stream));
  HIPCHECK(s, hipMemsetAsync(s->ctx.motion, 0, (4 * s->ctx.N + 12 * s->ctx.NF) * sizeof(float), s->stream));
  HIPCHECK(s, hipMemsetAsync(s->mg.vec, 0, 8 * s->plan.node_count * sizeof(float), s->stream));
  HIPCHECK(s, hipMemsetAsync(s->mg.snapshot, 0, 9 * s->ctx.N * sizeof(float), s->stream));
  HIPCHECK(s, hipMemsetAsync(s->mg.previous, 0, 9 * s->ctx.N * sizeof(float), s->stream));
  HIPCHECK(s, hipMemsetAsync(s->mg.state, 0, (64 + 32 * s->history_capacity) * sizeof(unsigned), s->stream));
  HIPCHECK(s, hipMemsetAsync(s->ctx.history, 0, (s->history_capacity * 4 + 4) * sizeof(float4), s->stream));
  HIPCHECK(s, hipMemsetAsync(s->ctx.chist, 0, (s->history_capacity * 5 + 5) * sizeof(float4), s->stream));
  HIPCHECK(s, hipMemsetAsync(s->ctx.tstate, 0, 8 * sizeof(unsigned long long), s->stream));
  HIPCHECK(s, hipMemsetAsync(s->ctx.step_times, 0, 4 * (s->history_capacity + 1) * sizeof(unsigned long long), s->stream));
  for (int op : {0, 40}) dispatch(s, op);
  if (s->ctx.pert_amp != 0) dispatch(s, 90);
  for (int op : {17, 18, 16, 19, 41, 42, 49, 85, 86}) dispatch(s, op);
  HIPCHECK(s, hipGetLastError());
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  // a reset is a new solver in the reference (water/lab_runner.gd:15-31): no
  // submission in flight, scheduler counters from zero
  s->sub_head = s->sub_count = 0;
  s->sched = wv_schedule_stats{};
  s->ema_batch_ms = 0;
  HIPCHECK(s, hipMemcpy(s->last_control, s->ctx.c, sizeof s->last_control, hipMemcpyDeviceToHost));
  return WV_OK;
}

extern "C" int wv_request_steps(wv_solver* s, int64_t steps) {
  if (!s || steps < 0) return WV_EINVAL;
  s->quota.store((double)steps);
  return WV_OK;
}
extern "C" int wv_request_time(wv_solver* s, double seconds) {
  if (!s || !std::isfinite(seconds) || seconds < 0) return WV_EINVAL;
  s->quota.store(-seconds);
  return WV_OK;
}

// Retire completed submissions, oldest first (consume_timestamps and the
// buffer_get_data_async callback of water/work_scheduler.gd:18-40, 80-87):
// the newest readback becomes the control copy, each submission's GPU time
// (events around its batches and readback) feeds the per-batch cost EMA
// (lerp 0.25 towards the new duration, as the reference's chunk costs) and
// the GPU-time total.  wait: 0 none, 1 every submission, 2 the oldest one.
static int retire(wv_solver* s, int wait, int* fresh) {
  if (wait == 1) HIPCHECK(s, hipStreamSynchronize(s->stream));
  bool first = true;
  while (s->sub_count > 0) {
    auto& b = s->sub[s->sub_head];
    if (wait == 2 && first) HIPCHECK(s, hipEventSynchronize(b.end));
    first = false;
    hipError_t e = hipEventQuery(b.end);
    if (e == hipErrorNotReady) break;
    if (e != hipSuccess) return fail(s, WV_EHIP, std::string("control readback: ") + hipGetErrorString(e));
    float ms = 0;
    HIPCHECK(s, hipEventElapsedTime(&ms, b.start, b.end));
    const double per = b.batches > 0 ? ms / b.batches : ms;
    s->ema_batch_ms = s->ema_batch_ms > 0 ? s->ema_batch_ms + 0.25 * (per - s->ema_batch_ms) : per;
    s->sched.gpu_ms += ms;
    s->sched.readbacks++;
    s->sched.inflight_batches -= b.batches;
    memcpy(s->last_control, s->ctrl_host + WV_NCTRL * s->sub_head, sizeof s->last_control);
    s->sub_head = (s->sub_head + 1) % WV_MAX_INFLIGHT;
    s->sub_count--;
    if (fresh) *fresh = 1;
  }
  s->sched.inflight = s->sub_count;
  s->sched.ema_batch_ms = s->ema_batch_ms;
  s->sched.accepted_time = s->last_control[1];
  s->sched.accepted_steps = s->last_control[2];
  s->sched.r_fluid = s->last_control[1] > 0 ? s->sched.gpu_ms * 1e-3 / s->last_control[1] : -1;
  return 0;
}

// frame() (water/work_scheduler.gd:41-87): one submission of up to
// max_batches whole batches, as many as the per-batch cost estimate fits in
// budget_ms (at least one), followed by one asynchronous control readback;
// nothing when WV_MAX_INFLIGHT submissions are unfinished.  The quota is
// uploaded inside the submission from its own pinned slot, so no earlier
// submission is waited for.
extern "C" int wv_advance(wv_solver* s, int max_batches, float budget_ms) {
  if (!s || max_batches < 0 || !(budget_ms >= 0)) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  int rc = retire(s, 0, nullptr); if (rc) return rc;
  if (max_batches == 0 || s->sub_count >= WV_MAX_INFLIGHT) return 0;
  auto t0 = std::chrono::steady_clock::now();
  int n = max_batches;
  if (budget_ms > 0 && s->ema_batch_ms > 0) n = std::max(1, std::min(n, (int)(budget_ms / s->ema_batch_ms)));
  const int slot = (s->sub_head + s->sub_count) % WV_MAX_INFLIGHT;
  auto& b = s->sub[slot];
  const float q = (float)s->quota.load();
  if (q != s->last_uploaded_quota) {
    // slot 1 + k belongs to submission slot k: reused only after that submission retired
    s->quota_host[1 + slot] = q;
    HIPCHECK(s, hipMemcpyAsync(s->quota_dev, s->quota_host + 1 + slot, sizeof(float), hipMemcpyHostToDevice, s->stream));
    s->last_uploaded_quota = q;
  }
  HIPCHECK(s, hipEventRecord(b.start, s->stream));
  if (s->use_graph) {
    rc = ensure_graph(s); if (rc) return rc;
    for (int k = 0; k < n; k++) HIPCHECK(s, hipGraphLaunch(s->batch_exec, s->stream));
  } else {
    for (int k = 0; k < n; k++) record_batch(s);
    HIPCHECK(s, hipGetLastError());
  }
  HIPCHECK(s, hipMemcpyAsync(s->ctrl_host + WV_NCTRL * slot, s->ctx.c, WV_NCTRL * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPCHECK(s, hipEventRecord(b.end, s->stream));
  b.batches = n;
  s->sub_count++;
  s->sched.submissions++;
  s->sched.batches += n;
  s->sched.inflight = s->sub_count;
  s->sched.inflight_batches += n;
  s->sched.max_inflight = std::max(s->sched.max_inflight, s->sub_count);
  s->sched.record_us += std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  return n;
}

extern "C" int wv_poll_control(wv_solver* s, float* out96, int* fresh, int wait) {
  if (!s || !out96 || wait < 0 || wait > 2) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  if (fresh) *fresh = 0;
  int rc = retire(s, wait, fresh); if (rc) return rc;
  memcpy(out96, s->last_control, sizeof s->last_control);
  return WV_OK;
}

extern "C" int wv_get_schedule_stats(wv_solver* s, wv_schedule_stats* out) {
  if (!s || !out) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  *out = s->sched;
  return WV_OK;
}

static void* field_ptr(wv_solver* s, int field, int64_t* bytes) {
  Ctx& c = s->ctx; MGDev& m = s->mg;
  switch (field) {
  case WV_FIELD_BANK_A: *bytes = c.stride * 4; return c.bank[0];
  case WV_FIELD_BANK_B: *bytes = c.stride * 4; return c.bank[1];
  case WV_FIELD_WORK: *bytes = 14 * c.N * 4; return c.work;
  case WV_FIELD_GEO: *bytes = 2 * c.N * 4; return c.geo;
  case WV_FIELD_CONTROL: *bytes = WV_NCTRL * 4; return c.c;
  case WV_FIELD_PLANE: *bytes = 4 * c.N * 4; return c.plane;
  case WV_FIELD_MOTION: *bytes = (4 * c.N + 12 * c.NF) * 4; return c.motion;
  case WV_FIELD_IMG_PLANES: *bytes = 4 * c.N * 4; return c.img_planes;
  case WV_FIELD_IMG_KINDS: *bytes = c.N * 4; return c.img_kinds;
  case WV_FIELD_HISTORY: *bytes = s->history_capacity * 4 * 16; return c.history;
  case WV_FIELD_COASTAL_HISTORY: *bytes = s->history_capacity * 5 * 16; return c.chist;
  case WV_FIELD_PROBE: *bytes = 8 * 16; return c.probe;
  case WV_FIELD_PARTIAL: *bytes = ((c.N + 127) / 128) * 16; return c.partial;
  case WV_FIELD_MG_VECTORS: *bytes = 8 * s->plan.node_count * 4; return m.vec;
  case WV_FIELD_SNAPSHOT: *bytes = 9 * c.N * 4; return m.snapshot;
  case WV_FIELD_AUDIT: *bytes = (64 + 32 * s->history_capacity) * 4; return m.state;
  case WV_FIELD_MG_ROWS: *bytes = s->plan.row_count * 4; return m.rows;
  case WV_FIELD_MG_LINV: { int64_t n = s->plan.lv[s->plan.L - 1].count; *bytes = n * n * 4; return m.linv; }
  case WV_FIELD_PRESSURE64: *bytes = c.N * 8; return c.pacc;
  case WV_FIELD_PREVIOUS: *bytes = 9 * c.N * 4; return m.previous;
  case WV_FIELD_STEP_TIMES: *bytes = s->history_capacity * 4 * 8; return c.step_times;
  case WV_FIELD_QUEUES: *bytes = 4 * 4; return c.mq_count;
  }
  *bytes = -1;
  return nullptr;
}

extern "C" int64_t wv_field_bytes(wv_solver* s, int field) {
  if (!s) return WV_EINVAL;
  int64_t b; field_ptr(s, field, &b); return b;
}
extern "C" int wv_read_field(wv_solver* s, int field, void* dst, int64_t bytes) {
  if (!s || !dst) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  int64_t n; void* p = field_ptr(s, field, &n);
  if (!p || bytes < 0 || bytes > n) return fail(s, WV_EINVAL, "bad field or size");
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  HIPCHECK(s, hipMemcpy(dst, p, (size_t)bytes, hipMemcpyDeviceToHost));
  return WV_OK;
}
extern "C" int wv_write_field(wv_solver* s, int field, const void* src, int64_t bytes) {
  if (!s || !src) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  int64_t n; void* p = field_ptr(s, field, &n);
  if (!p || bytes < 0 || bytes > n || field == WV_FIELD_QUEUES) return fail(s, WV_EINVAL, "bad field or size");
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  HIPCHECK(s, hipMemcpy(p, src, (size_t)bytes, hipMemcpyHostToDevice));
  return WV_OK;
}

extern "C" int wv_op(wv_solver* s, int op) {
  if (!s) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  int rc = sync_quota(s); if (rc) return rc;
  rc = dispatch(s, op); if (rc) return rc;
  HIPCHECK(s, hipGetLastError());
  return WV_OK;
}

extern "C" int wv_mg_record(wv_solver* s, int build, int initial) {
  if (!s) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  if (build) wv_mg_build(s->ctx, s->mg, s->stream);
  else wv_mg_cycle(s->ctx, s->mg, initial ? 0 : 1, s->smoothing, true, s->stream);
  HIPCHECK(s, hipGetLastError());
  return WV_OK;
}

// ---- state file (no reference counterpart: its buffers are GPU-only and
// created cleared, water/pressure_gpu.gd:93-100) ------------------------------
// Layout, little endian: a 256-byte header (magic "WAVESTAT", version, field
// count, grid, the 40 setting floats), then one 32-byte table entry per field
// (id, bytes per element, bytes, file offset, checksum), then each
// field's bytes at a 256-byte aligned offset.
namespace {
const char kMagic[8] = {'W', 'A', 'V', 'E', 'S', 'T', 'A', 'T'};
const int kStateFields[] = {WV_FIELD_BANK_A, WV_FIELD_BANK_B, WV_FIELD_WORK, WV_FIELD_GEO, WV_FIELD_CONTROL,
                            WV_FIELD_PLANE, WV_FIELD_MOTION, WV_FIELD_IMG_PLANES, WV_FIELD_IMG_KINDS,
                            WV_FIELD_HISTORY, WV_FIELD_COASTAL_HISTORY, WV_FIELD_PROBE, WV_FIELD_MG_VECTORS,
                            WV_FIELD_SNAPSHOT, WV_FIELD_AUDIT, WV_FIELD_MG_ROWS, WV_FIELD_MG_LINV,
                            WV_FIELD_PREVIOUS, WV_FIELD_STEP_TIMES};
struct StateHeader {
  char magic[8];
  uint32_t version, field_count;
  int32_t n[3];
  uint32_t header_bytes;
  float pc[32], opts[8];
  uint8_t pad[256 - 8 - 8 - 12 - 4 - 160];
};
static_assert(sizeof(StateHeader) == 256, "state header is 256 bytes");
struct StateEntry { uint32_t id, elem; uint64_t bytes, offset, checksum; };
static_assert(sizeof(StateEntry) == 32, "state table entry is 32 bytes");
// checksum: sum over the little-endian 32-bit words w_i (the last one
// zero-padded) of (w_i + 1) (2 i + 1) K, modulo 2^64, K = 0x9E3779B97F4A7C15
uint64_t state_sum(const unsigned char* p, size_t n) {
  const uint64_t K = 0x9E3779B97F4A7C15ull;
  uint64_t h = 0;
  for (size_t i = 0; 4 * i < n; i++) {
    uint32_t w = 0;
    memcpy(&w, p + 4 * i, n - 4 * i >= 4 ? 4 : n - 4 * i);
    h += ((uint64_t)w + 1) * ((2 * (uint64_t)i + 1) * K);
  }
  return h;
}
int field_elem(int id) { return id == WV_FIELD_STEP_TIMES ? 8 : 4; }
// settings that define the buffers and the physics; the op and quota slots are
// per dispatch, and the PCG iteration slots per batch (opts[5]) only split the
// same gated program differently
bool same_settings(const wv_settings& a, const float* pc, const float* opts) {
  for (int k = 0; k < 32; k++) if (k != 3 && k != 15 && memcmp(&a.pc[k], &pc[k], 4) != 0) return false;
  for (int k = 0; k < 8; k++) if (k != 5 && memcmp(&a.opts[k], &opts[k], 4) != 0) return false;
  return true;
}
}  // namespace

extern "C" int wv_save_state(wv_solver* s, const char* path) {
  if (!s || !path) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  const int nf = (int)(sizeof kStateFields / sizeof kStateFields[0]);
  StateHeader h{};
  memcpy(h.magic, kMagic, 8);
  h.version = WV_STATE_VERSION; h.field_count = nf;
  h.n[0] = s->ctx.nx; h.n[1] = s->ctx.ny; h.n[2] = s->ctx.nz;
  h.header_bytes = sizeof h;
  memcpy(h.pc, s->settings.pc, sizeof h.pc); memcpy(h.opts, s->settings.opts, sizeof h.opts);
  std::vector<StateEntry> table(nf);
  uint64_t at = (sizeof h + nf * sizeof(StateEntry) + 255) / 256 * 256;
  std::vector<std::vector<unsigned char>> data(nf);
  for (int k = 0; k < nf; k++) {
    int64_t bytes; void* p = field_ptr(s, kStateFields[k], &bytes);
    if (!p || bytes < 0) return fail(s, WV_EINVAL, "state field missing");
    data[k].resize((size_t)bytes);
    if (bytes) HIPCHECK(s, hipMemcpy(data[k].data(), p, (size_t)bytes, hipMemcpyDeviceToHost));
    table[k] = {(uint32_t)kStateFields[k], (uint32_t)field_elem(kStateFields[k]), (uint64_t)bytes, at,
                state_sum(data[k].data(), data[k].size())};
    at = (at + (uint64_t)bytes + 255) / 256 * 256;
  }
  FILE* f = fopen(path, "wb");
  if (!f) return fail(s, WV_EINVAL, std::string("cannot write ") + path);
  bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(table.data(), sizeof(StateEntry), nf, f) == (size_t)nf;
  for (int k = 0; ok && k < nf; k++)
    ok = fseek(f, (long)table[k].offset, SEEK_SET) == 0 && (data[k].empty() || fwrite(data[k].data(), 1, data[k].size(), f) == data[k].size());
  ok = fclose(f) == 0 && ok;
  return ok ? WV_OK : fail(s, WV_EINVAL, std::string("short write to ") + path);
}

extern "C" int wv_load_state(wv_solver* s, const char* path) {
  if (!s || !path) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  FILE* f = fopen(path, "rb");
  if (!f) return fail(s, WV_EINVAL, std::string("cannot read ") + path);
  struct Closer { FILE* f; ~Closer() { fclose(f); } } closer{f};
  StateHeader h;
  if (fread(&h, sizeof h, 1, f) != 1 || memcmp(h.magic, kMagic, 8) != 0) return fail(s, WV_EINVAL, "not a state file");
  if (h.version != WV_STATE_VERSION) return fail(s, WV_EINVAL, "state file version " + std::to_string(h.version));
  if (h.n[0] != s->ctx.nx || h.n[1] != s->ctx.ny || h.n[2] != s->ctx.nz || !same_settings(s->settings, h.pc, h.opts))
    return fail(s, WV_EINVAL, "state file settings differ from the solver's");
  if (h.field_count > 64) return fail(s, WV_EINVAL, "corrupt field table");
  std::vector<StateEntry> table(h.field_count);
  if (fread(table.data(), sizeof(StateEntry), h.field_count, f) != h.field_count) return fail(s, WV_EINVAL, "short field table");
  // read and check every field before anything reaches the device
  std::vector<std::vector<unsigned char>> data(h.field_count);
  for (uint32_t k = 0; k < h.field_count; k++) {
    int64_t bytes; void* p = field_ptr(s, (int)table[k].id, &bytes);
    if (!p || (uint64_t)bytes != table[k].bytes) return fail(s, WV_EINVAL, "field " + std::to_string(table[k].id) + " does not fit");
    data[k].resize((size_t)bytes);
    if (fseek(f, (long)table[k].offset, SEEK_SET) != 0 || (bytes && fread(data[k].data(), 1, (size_t)bytes, f) != (size_t)bytes))
      return fail(s, WV_EINVAL, "short field " + std::to_string(table[k].id));
    if (state_sum(data[k].data(), data[k].size()) != table[k].checksum) return fail(s, WV_EINVAL, "checksum of field " + std::to_string(table[k].id));
  }
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  for (uint32_t k = 0; k < h.field_count; k++) {
    int64_t bytes; void* p = field_ptr(s, (int)table[k].id, &bytes);
    if (bytes) HIPCHECK(s, hipMemcpy(p, data[k].data(), (size_t)bytes, hipMemcpyHostToDevice));
  }
  // the control copy and the pressure-timing state follow the loaded vector
  HIPCHECK(s, hipMemcpy(s->last_control, s->ctx.c, sizeof s->last_control, hipMemcpyDeviceToHost));
  unsigned long long t[8] = {0, 0, 0, 0, (unsigned long long)s->last_control[2], (unsigned long long)s->last_control[32], 0, 0};
  HIPCHECK(s, hipMemcpy(s->ctx.tstate, t, sizeof t, hipMemcpyHostToDevice));
  return WV_OK;
}

extern "C" int wv_synchronize(wv_solver* s) {
  if (!s) return WV_EINVAL;
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  return WV_OK;
}

extern "C" int wv_plan_levels(const wv_settings* settings, int* meta16, int max_levels) {
  if (!settings || max_levels < 0) return WV_EINVAL;
  const float* pc = settings->pc;
  if (pc[0] < 1 || pc[1] < 1 || pc[2] < 1 || pc[4] <= 0 || pc[5] <= 0 || pc[6] <= 0) return WV_EINVAL;
  Plan p;
  if (!plan_levels(pc, (int)settings->opts[3], p)) return WV_EPLAN;
  if (meta16) for (int l = 0; l < p.L && l < max_levels; l++) memcpy(meta16 + 16 * l, &p.meta[16 * l], 64);
  return p.L;
}

extern "C" int wv_levels(wv_solver* s, int* meta16, int max_levels) {
  if (!s) return WV_EINVAL;
  int L = s->plan.L;
  if (meta16) for (int l = 0; l < L && l < max_levels; l++) memcpy(meta16 + 16 * l, &s->plan.meta[16 * l], 64);
  return L;
}

void wv_launch_timed(const Ctx& c, const MGDev& m, int kernel, hipStream_t st);
int wv_time_copy(float* a, float* b, int64_t n, int reps, hipStream_t st, float* ms);
extern "C" int wv_time_kernel(wv_solver* s, int kernel, int reps, float* avg_us) {
  if (!s || !avg_us || reps < 1 || kernel < 0 || kernel > 6) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  Ctx c = s->ctx;
  if (kernel == 6) {
    // plain copy of 2^28 floats (1 GiB read + 1 GiB written): a working set
    // far beyond the 256 MB last-level cache, the achievable-bandwidth reference
    const int64_t n = (int64_t)1 << 28;
    float *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 4) != hipSuccess) return fail(s, WV_ENOMEM, "copy reference: 1 GiB");
    if (hipMalloc(&b, n * 4) != hipSuccess) { hipFree(a); return fail(s, WV_ENOMEM, "copy reference: 1 GiB"); }
    float ms = 0;
    int rc = wv_time_copy(a, b, n, reps, s->stream, &ms);
    hipFree(a); hipFree(b);
    if (rc) return fail(s, WV_EHIP, "copy reference failed");
    *avg_us = ms * 1000.0f / reps;
    return WV_OK;
  }
  if (kernel >= 4) {
    // the step kernels gate themselves on the control vector: time them against
    // a copy of it with a trial open (c[4] = c[5] = 1, no convergence, no failure)
    float host[WV_NCTRL];
    HIPCHECK(s, hipStreamSynchronize(s->stream));
    HIPCHECK(s, hipMemcpy(host, s->ctx.c, sizeof host, hipMemcpyDeviceToHost));
    host[4] = 1; host[5] = 1; host[6] = 0; host[7] = 0; host[57] = 0;
    HIPCHECK(s, hipMemcpy(s->ctl_timing, host, sizeof host, hipMemcpyHostToDevice));
    c.c = s->ctl_timing;
  }
  auto launch = [&]() {
    if (kernel == 5) wv_launch_op(c, 27, s->stream);   // MacCormack forward trace
    else wv_launch_timed(c, s->mg, kernel, s->stream);
  };
  hipEvent_t a, b;
  HIPCHECK(s, hipEventCreate(&a)); HIPCHECK(s, hipEventCreate(&b));
  launch();  // warm
  HIPCHECK(s, hipEventRecord(a, s->stream));
  for (int r = 0; r < reps; r++) launch();
  HIPCHECK(s, hipEventRecord(b, s->stream));
  HIPCHECK(s, hipEventSynchronize(b));
  float ms = 0; HIPCHECK(s, hipEventElapsedTime(&ms, a, b));
  hipEventDestroy(a); hipEventDestroy(b);
  *avg_us = ms * 1000.0f / reps;
  return WV_OK;
}

// Standalone pressure solve on the current geometry (see include/waves.h).
extern "C" int wv_pressure_solve(wv_solver* s, int mode, float tolerance, int max_iterations, int flags, wv_solve_report* out) {
  if (!s || !out || mode < 0 || mode > 2 || !(tolerance > 0) || max_iterations < 1) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  if (mode == 2 && !s->net) return fail(s, WV_EINVAL, "mode 2 needs a network (wv_solver_attach_net)");
  const bool zero_mean = (flags & WV_SOLVE_ZERO_MEAN) != 0, given = (flags & WV_SOLVE_GIVEN_RHS) != 0;
  Ctx c = s->ctx; c.tol = tolerance; c.flex = mode == 2;
  const MGDev& m = s->mg; hipStream_t st = s->stream;
  // finish and retire every submission first: the solve reads back through the
  // first readback slot
  int rc = retire(s, 1, nullptr); if (rc) return rc;
  struct Events {  // destroyed on every return path
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() { if (a) hipEventDestroy(a); if (b) hipEventDestroy(b); }
  } ev;
  HIPCHECK(s, hipEventCreate(&ev.a)); HIPCHECK(s, hipEventCreate(&ev.b));
  hipEvent_t e0 = ev.a, e1 = ev.b;
  // the solve opens its own trial in the control vector; the caller's control
  // vector is restored afterwards (the accepted bank is never written)
  HIPCHECK(s, hipMemcpyAsync(s->ctl_saved, c.c, WV_NCTRL * sizeof(float), hipMemcpyDeviceToDevice, st));
  auto wall0 = std::chrono::steady_clock::now();
  HIPCHECK(s, hipEventRecord(e0, st));
  wv_solve_begin(c, !given, st);
  wv_mg_build(c, m, st);                       // hierarchy audit + Galerkin (skipped if unchanged)
  if (mode >= 1) wv_mg_to_bf16(c, m, st);
  if (given) wv_solve_init_given(c, st);       // r = b - A p from the caller's b and p
  else {
    wv_launch_op(c, 3, st);                    // b, r = b - A p, z, p-dir
    if (zero_mean) wv_solve_zero_mean(c, st);
  }
  if (mode == 2) wv_solve_learned_start(c, s->net, st);
  auto cycle = [&](int phase, bool load) {
    if (mode >= 1) wv_mg_cycle_bf16(c, m, phase, s->smoothing, load, true, st);
    else wv_mg_cycle(c, m, phase, s->smoothing, load, st);
  };
  cycle(0, true);                              // z = M r, rho, convergence (op 4)
  int restarts = 0;
  const char* trace = getenv("WAVES_SOLVE_TRACE");
  const bool tracing = trace && trace[0] == '1';
  // iterations run in chunks, each followed by one control readback; the
  // first chunk is as long as the previous solve of this form needed, later
  // ones two iterations.  Iterations past convergence are predicated off on
  // the device.  (Capturing a chunk into a hipGraph and replaying it was
  // measured 7 % slower than these direct launches.)
  auto iterate = [&](int n) {
    for (int k = 0; k < n; k++) {
      if (mode >= 1) { wv_launch_op(c, 5, st); wv_solve_dot_alpha64(c, st); }   // A p, alpha in FP64
      else wv_launch_spmv_dot(c, st);          // A p with the p.Ap partials, alpha
      wv_launch_pcg_update(c, m, st);
      cycle(1, false);                         // z = M r, beta, convergence (op 9)
      wv_launch_op(c, 10, st);                 // p = z + beta p
    }
    return hipMemcpyAsync(s->ctrl_host, c.c, WV_NCTRL * sizeof(float), hipMemcpyDeviceToHost, st);
  };
  int chunk = tracing ? 1 : std::max(1, std::min({16, max_iterations, s->solve_iterations[mode] > 0 ? s->solve_iterations[mode] : 4}));
  float ctl[WV_NCTRL];
  for (;;) {
    HIPCHECK(s, iterate(chunk));
    HIPCHECK(s, hipStreamSynchronize(st));
    chunk = tracing ? 1 : 2;
    memcpy(ctl, s->ctrl_host, sizeof ctl);
    if (chunk == 1) {
      double d[8];
      HIPCHECK(s, hipMemcpyAsync(d, c.dctl, sizeof d, hipMemcpyDeviceToHost, st));
      HIPCHECK(s, hipStreamSynchronize(st));
      fprintf(stderr, "[solve mode %d] it %g conv %g fail %g relres %g alpha %g beta %g rho %g b2 %g r2 %g dt %g tol %g\n",
              mode, ctl[8], ctl[6], ctl[7], ctl[13], ctl[11], ctl[12], d[0], d[1], d[2], ctl[20], c.tol);
    }
    if (ctl[7] > .5f || ctl[8] >= max_iterations) break;
    if (ctl[6] > .5f) {
      if (mode == 0) break;
      wv_solve_true_residual64(c, m, true, 3, st);  // confirm in FP64, else fold p, replace r, restart
      double flag = 0;
      HIPCHECK(s, hipMemcpyAsync(&flag, c.dctl + 6, sizeof flag, hipMemcpyDeviceToHost, st));
      HIPCHECK(s, hipStreamSynchronize(st));
      if (flag < .5) break;
      if (++restarts > 16) break;
      cycle(0, true);
    }
  }
  wv_solve_finish(c, st);                      // FP64 solution; candidate p = its FP32 rounding
  wv_launch_op(c, 11, st);                     // projection (pressure_pass.glslinc:52-62)
  HIPCHECK(s, hipEventRecord(e1, st));
  HIPCHECK(s, hipEventSynchronize(e1));
  auto wall1 = std::chrono::steady_clock::now();
  float ms = 0; HIPCHECK(s, hipEventElapsedTime(&ms, e0, e1));
  wv_solve_true_residual64(c, m, false, 2, st);   // report on the FP64 solution
  double d[16];
  HIPCHECK(s, hipMemcpyAsync(d, c.dctl, sizeof d, hipMemcpyDeviceToHost, st));
  HIPCHECK(s, hipMemcpyAsync(s->ctrl_host, c.c, sizeof ctl, hipMemcpyDeviceToHost, st));
  HIPCHECK(s, hipStreamSynchronize(st));
  memcpy(ctl, s->ctrl_host, sizeof ctl);
  HIPCHECK(s, hipMemcpyAsync(c.c, s->ctl_saved, WV_NCTRL * sizeof(float), hipMemcpyDeviceToDevice, st));
  HIPCHECK(s, hipStreamSynchronize(st));
  s->solve_iterations[mode] = (int)ctl[8];
  out->iterations = (int)ctl[8];
  out->converged = ctl[6] > .5f;
  out->failure = (int)ctl[7];
  out->restarts = restarts;
  out->device_ms = ms;
  out->wall_ms = (float)std::chrono::duration<double, std::milli>(wall1 - wall0).count();
  out->true_rel_residual = d[4] > 0 ? std::sqrt(d[3] / d[4]) : 0;
  out->recurrence_rel_residual = ctl[13];
  return WV_OK;
}

extern "C" int wv_solver_attach_net(wv_solver* s, wv_net* net) {
  if (!s) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  if (net) {
    int d[4]; wv_net_grid(net, d);
    if (d[0] != s->ctx.nx || d[1] != s->ctx.ny || d[2] != s->ctx.nz || d[3] != 1)
      return fail(s, WV_EINVAL, "network grid differs from the solver grid or batch != 1");
  }
  s->net = net;
  return WV_OK;
}

extern "C" int wv_set_precision(wv_solver* s, int mode) {
  if (!s || mode < 0 || mode > 1) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  if (mode == s->precision) return WV_OK;
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  if (s->batch_exec) { HIPCHECK(s, hipGraphExecDestroy(s->batch_exec)); s->batch_exec = nullptr; }
  if (s->batch_graph) { HIPCHECK(s, hipGraphDestroy(s->batch_graph)); s->batch_graph = nullptr; }
  s->precision = mode;
  return WV_OK;
}

extern "C" int wv_liquid_volume(wv_solver* s, int bank, double* volume) {
  if (!s || !volume || (bank != 0 && bank != 1)) return WV_EINVAL;
  std::lock_guard<std::mutex> g(s->lock);
  wv_solve_volume64(s->ctx, bank, s->stream);
  double d[16];
  HIPCHECK(s, hipMemcpyAsync(d, s->ctx.dctl, sizeof d, hipMemcpyDeviceToHost, s->stream));
  HIPCHECK(s, hipStreamSynchronize(s->stream));
  *volume = d[13];
  return WV_OK;
}
