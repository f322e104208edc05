/*
 * waves_hip -- C-ABI of the MI355X (gfx950) free-surface solver step.
 *
 * Drop-in for the GPU side of the reference's solver runtime: the Vulkan
 * compute dispatch of water/work_scheduler.gd, the kernels under
 * water/shaders/compute, and the multigrid hierarchy of
 * water/pressure_multigrid.gd and water/pressure_hierarchy_state.gd.  The host
 * keeps the reference's step / reset / readback surface
 * (water/pressure_gpu.gd:40-59, 133-137) and its staggered MAC buffer layout.
 *
 * Conventions: plain pointers and sizes, no C++ or framework types; every
 * function returns 0 on success and a negative WV_E* code on a host or HIP
 * error (never throws); wv_last_error() returns the message.  Device failure
 * codes stay in the control vector c[7] exactly as in the reference
 * (pressure_pass.glslinc, coupled_motion.glslinc, coastal_pass.glslinc).
 * One solver handle owns its device memory and one HIP stream.  request_*
 * may be called from any thread; all other calls from one thread per handle.
 */
#ifndef WAVES_HIP_H
#define WAVES_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct wv_solver wv_solver;

/* Settings: the 32-float parameter block of water/pressure_gpu.gd:68 +
 * water/coupled_gpu.gd:18-21 + water/coastal_gpu.gd:23-24 (index 3 = op and
 * 15 = quota are ignored), plus 8 option floats: right absorption
 * (water/coastal_gpu.gd:12), MG smoothing / lag / aggregate_from / maintenance
 * mode (water/pressure_gpu.gd:42-43), PCG iterations per batch
 * (water/coupled_gpu.gd:14), alongshore perturbation amplitude and seed. */
typedef struct wv_settings {
  float pc[32];
  float opts[8];
} wv_settings;

enum {
  WV_OK = 0,
  WV_EINVAL = -1,   /* bad argument or size */
  WV_EHIP = -2,     /* a HIP runtime call failed */
  WV_ENOMEM = -3,   /* device allocation failed */
  WV_ENODEV = -4,   /* no gfx950 device */
  WV_EPLAN = -5     /* multigrid planning failed (reference host failure 97) */
};

/* Field ids for wv_read_field / wv_write_field (layouts of
 * pressure_core.glslinc:2-11 and coupled_motion.glslinc:1-14). */
enum {
  WV_FIELD_BANK_A = 0,        /* [F N][U][V][W][P N] float */
  WV_FIELD_BANK_B = 1,
  WV_FIELD_WORK = 2,          /* 14 N float */
  WV_FIELD_GEO = 3,           /* 2 N float (phi, kind) */
  WV_FIELD_CONTROL = 4,       /* 96 float, the reference's c[] */
  WV_FIELD_PLANE = 5,         /* 4 N float, reconstructed local PLIC */
  WV_FIELD_MOTION = 6,        /* 4 N + 12 faces float */
  WV_FIELD_IMG_PLANES = 7,    /* 4 N float, exported accepted planes */
  WV_FIELD_IMG_KINDS = 8,     /* N float, exported kinds 0/1/2 */
  WV_FIELD_HISTORY = 9,       /* capacity*4 float4 step history */
  WV_FIELD_COASTAL_HISTORY = 10, /* capacity*5 float4 */
  WV_FIELD_PROBE = 11,        /* 8 float4 analytic probes */
  WV_FIELD_PARTIAL = 12,      /* ceil(N/128) float4 */
  WV_FIELD_MG_VECTORS = 13,   /* 8 floats per MG node, per level */
  WV_FIELD_SNAPSHOT = 14,     /* 9 N float, preconditioner snapshot */
  WV_FIELD_AUDIT = 15,        /* hierarchy audit state (uint32) */
  WV_FIELD_MG_ROWS = 16,      /* coarse operator rows, k-major, levels >= 1 */
  WV_FIELD_MG_LINV = 17,      /* last level: dense L^-1, row-major */
  WV_FIELD_PRESSURE64 = 18,   /* N double: pressure of the last standalone solve in FP64 */
  WV_FIELD_PREVIOUS = 19,     /* 9 N float, the hierarchy audit's previous-trial operator copy */
  WV_FIELD_STEP_TIMES = 20,   /* capacity x 4 uint64 per accepted step k (row k - 1): device ns of the
                                 accepted trial's pressure sections (hierarchy audit/rebuild, PCG, projection,
                                 defect corrections), device ns of every trial's pressure sections since
                                 the previous acceptance, the trial count c[32], the sections timed */
  WV_FIELD_QUEUES = 21        /* 4 uint32, read only: faces or cells the last queued passes took on the
                                 general path (MacCormack forward trace, MacCormack correction, exact distance) */
};

/* _init(settings) + start() + _initialize (water/pressure_gpu.gd:40-47, 82-127):
 * plans the MG hierarchy (water/pressure_multigrid.gd:33-134), allocates every
 * device buffer, clears them. Does not run the initial commands. */
int wv_create(const wv_settings *settings, wv_solver **out);
/* close() / _free() (water/pressure_gpu.gd:55-59, 138-149) */
int wv_destroy(wv_solver *s);
/* initial command list (water/coupled_gpu.gd:34-37, water/coastal_gpu.gd:26-27);
 * a reset of the reference is close + new + start (water/lab_runner.gd:15-31). */
int wv_reset(wv_solver *s);
/* request(steps) / request_time(seconds) (water/pressure_gpu.gd:48-54): the
 * quota the device's trial start checks (pressure_pass.glslinc:20). Thread-safe. */
int wv_request_steps(wv_solver *s, int64_t steps);
int wv_request_time(wv_solver *s, double seconds);
/* frame() (water/work_scheduler.gd:41-87): one non-blocking submission of
 * whole batches of the coupled step program (water/coupled_gpu.gd:38-56,
 * water/coastal_gpu.gd:28-33) with the current quota, followed by one
 * asynchronous 384-byte control readback.  At most max_batches; with
 * budget_ms > 0 only as many as the per-batch GPU cost estimate (an EMA of
 * measured submissions, as the reference's per-chunk costs) fits in the
 * budget, at least one (the reference: gpu_budget_ms 14); nothing while
 * WV_MAX_INFLIGHT submissions are unfinished (max_inflight_submissions 3).
 * Returns the number of batches enqueued (>= 0) or a negative WV_E* code. */
#define WV_MAX_INFLIGHT 3
int wv_advance(wv_solver *s, int max_batches, float budget_ms);
/* buffer_get_data_async(control) callback (water/work_scheduler.gd:80-87):
 * retires finished submissions in order and copies the newest readback into
 * out[96]; *fresh = 1 if a readback arrived since the last poll.  wait: 0 does
 * not block, 1 blocks until all enqueued work is done, 2 blocks until the
 * oldest unfinished submission is done. */
int wv_poll_control(wv_solver *s, float *out96, int *fresh, int wait);
/* Scheduler telemetry (water/work_scheduler.gd:18-40: gpu_ns, record_us,
 * R_fluid, queue depth).  Counters from the last wv_reset. */
typedef struct wv_schedule_stats {
  int64_t submissions;      /* wv_advance calls that enqueued work, one control readback each */
  int64_t readbacks;        /* readbacks retired by wv_poll_control / wv_advance */
  int64_t batches;          /* batches enqueued */
  int inflight;             /* submissions not yet retired */
  int max_inflight;         /* deepest queue seen (maximum_queue) */
  int64_t inflight_batches; /* batches of the unretired submissions */
  double gpu_ms;            /* device time of the retired submissions (total_gpu_ns / 1e6) */
  double record_us;         /* host time spent enqueueing (total_record_us) */
  double ema_batch_ms;      /* per-batch cost estimate */
  double accepted_time;     /* c[1] of the newest readback */
  double accepted_steps;    /* c[2] of the newest readback */
  double r_fluid;           /* GPU seconds per accepted simulated second, -1 before time advances */
} wv_schedule_stats;
int wv_get_schedule_stats(wv_solver *s, wv_schedule_stats *out);
/* parity / render export I/O (no reference counterpart: the reference's
 * buffers are GPU-only; the planes/kinds export replaces the Texture3DRD
 * upload of water/pressure_gpu.gd:106-115). Synchronous. */
int64_t wv_field_bytes(wv_solver *s, int field);
int wv_read_field(wv_solver *s, int field, void *dst, int64_t bytes);
int wv_write_field(wv_solver *s, int field, const void *src, int64_t bytes);
/* State file (no reference counterpart: the reference's buffers are GPU-only
 * and start cleared, water/pressure_gpu.gd:93-100; this is the identical-input
 * harness's format).  Version WV_STATE_VERSION, little endian: a 256-byte
 * header (magic "WAVESTAT", uint32 version, uint32 field count, int32 n[3],
 * uint32 header bytes, the 32 + 8 setting floats of wv_settings), one 32-byte
 * entry per field (uint32 field id, uint32 bytes per element, uint64 bytes,
 * uint64 file offset, uint64 checksum: the sum over the bytes' little-endian
 * 32-bit words w_i, the last zero-padded, of (w_i + 1)(2i + 1) 0x9E3779B97F4A7C15
 * modulo 2^64), then each
 * field's bytes at a 256-byte aligned offset.  The fields: both banks, work,
 * geo, control, plane, motion, exported planes and kinds, both history rings,
 * probes, multigrid vectors, snapshot, previous copy, audit state, coarse
 * rows, coarse L^-1 and the per-step pressure timing -- everything a batch
 * reads, so a loaded solver continues bit-identically.  wv_load_state checks
 * magic, version, grid, settings (except the per-dispatch op and quota slots),
 * sizes and checksums before anything reaches the device.  Both wait for the
 * handle's stream. */
#define WV_STATE_VERSION 1
int wv_save_state(wv_solver *s, const char *path);
int wv_load_state(wv_solver *s, const char *path);
/* one reference op (water/pressure_gpu.gd:69-81 _dispatch without the MG hook),
 * for the parity harness; ops whose finaliser is fused into the previous op's
 * kernel are no-ops. Enqueued, not waited for.  WV_OP_SPMV_DOT (no reference
 * counterpart) runs ops 5, 6 and 7 in the one-pass form the plain-FP32 batch
 * and standalone solve use. */
#define WV_OP_SPMV_DOT 105
int wv_op(wv_solver *s, int op);
/* multigrid.record(list, build, initial) (water/pressure_multigrid.gd:155-176) */
int wv_mg_record(wv_solver *s, int build, int initial);
/* wait for everything enqueued on the handle's stream */
int wv_synchronize(wv_solver *s);
const char *wv_last_error(wv_solver *s);
/* MG level descriptors: 16 ints per level, the reference's `meta` layout
 * (water/pressure_multigrid.gd:54, 61-68). Returns the level count. */
int wv_levels(wv_solver *s, int *meta16, int max_levels);
/* The same level plan computed on the host alone, without a device or a
 * handle: multigrid.initialize (water/pressure_multigrid.gd:33-69) is CPU
 * planning in the reference too.  Returns the level count, WV_EPLAN when the
 * grid cannot be planned. */
int wv_plan_levels(const wv_settings *settings, int *meta16, int max_levels);
/* Average device duration [us] of one launch of a hot kernel, timed with HIP
 * events on the solver's stream over `reps` launches on the current state
 * (kernel: 0 = level-0 Chebyshev sweep, 1 = fine masked 7-point SpMV,
 * 2 = level-0 residual, 3 = plain copy of N floats, the bandwidth reference,
 * 4 = Chebyshev sweep on MG level 1, 5 = MacCormack forward trace, op 27;
 * kernels 4 and 5 run against a copy of the control vector with a trial open;
 * 6 = streaming copy of 1 GiB into another 1 GiB, two buffers allocated for
 * the measurement and freed after it: the large-working-set bandwidth reference).
 * Writes only scratch that every trial recomputes before use (V-cycle / PCG /
 * Galerkin vectors, the MacCormack forward field and trace cache): the
 * accepted state is unchanged. */
int wv_time_kernel(wv_solver *s, int kernel, int reps, float *avg_us);

/* Report of one standalone pressure solve. */
typedef struct wv_solve_report {
  int iterations;                  /* PCG iterations (c[8]) */
  int converged;                   /* c[6] at the end */
  int failure;                     /* c[7] (1 breakdown, 2 NaN, 3 > 4096 iterations) */
  int restarts;                    /* mixed form: residual replacements */
  float device_ms;                 /* HIP-event time, residual initialisation .. projection */
  float wall_ms;                   /* host wall time of the same span */
  double true_rel_residual;        /* |b - A p| / |b| recomputed in FP64 on the device */
  float recurrence_rel_residual;   /* c[13] */
} wv_solve_report;

/* Standalone pressure solve on the current state (no reference counterpart:
 * the reference only solves inside its batch program). Opens a trial, rebuilds
 * the hierarchy when the audit asks for it, sets p = 0, forms b from the
 * candidate face velocities (op 3, optionally shifted to zero mean over the
 * wet cells), iterates PCG to `tolerance` with the reference's convergence
 * predicates, and projects (op 11).  mode 0: plain FP32 MGPCG (the reference
 * ops); mode 1: BF16 V-cycle operators, FP32 vectors, FP64 dot products and an
 * FP64 true-residual confirmation; a converged flag the true residual does not
 * confirm folds p into an FP64 accumulator, replaces r by the true residual
 * and restarts CG on the FP32 correction.  The FP64 solution is left in
 * WV_FIELD_PRESSURE64, its FP32 rounding in the candidate bank's p, and
 * out->true_rel_residual is computed from the FP64 solution.  mode 2: mode 1
 * with the learned warm start of the attached network (waves_net.h): the
 * network's two output fields span a correction of the starting pressure
 * chosen by an FP64 Galerkin projection (it cannot increase the A-norm
 * error), and CG uses the flexible (Polak-Ribiere) coefficient. */
/* Liquid volume sum_i F_i hx hy hz of bank 0 or 1, reduced in FP64 on the
 * device in a fixed order (no reference counterpart: the reference's volume,
 * c[17], is an FP32 sum).  Waits for the handle's stream. */
int wv_liquid_volume(wv_solver *s, int bank, double *volume);

/* Allocation-free variant of wv_liquid_volume: the same FP64 reduction (same
 * 128-lane groups, same order, bit-identical result) of `cells` volume
 * fractions at the device address `fraction`, times hx hy hz, written to the
 * device address `volume` (one double), enqueued on `stream` (a hipStream_t;
 * NULL: the null stream) and not waited for.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no allocating routine: its per-group partials live in a static
 * device array of WV_NOALLOC_MAX_CELLS / 128 doubles (1 MiB) that each GPU
 * holds until the process ends; larger inputs return WV_EINVAL.  Calls that
 * overlap on one GPU share that array and race. */
#define WV_NOALLOC_MAX_CELLS (1 << 24)
int wv_liquid_volume_noalloc(const float *fraction, int64_t cells, float hx, float hy, float hz, double *volume,
                             void *stream);

/* Precision of the batch program's pressure solve (no reference counterpart:
 * the reference always runs mode 0).  0: the reference's FP32 PCG and
 * V-cycle; 1: the V-cycle's operators in BF16 (FP32 vectors and accumulation)
 * and every PCG dot product and scalar in FP64, the reference's
 * post-projection check and defect correction unchanged.  Takes effect at the
 * next wv_advance. */
int wv_set_precision(wv_solver *s, int mode);

enum {
  WV_SOLVE_ZERO_MEAN = 1,  /* shift b to zero mean over the wet cells (starts from p = 0) */
  WV_SOLVE_GIVEN_RHS = 2   /* b is work field 5 as written by the caller; start from the candidate p */
};
int wv_pressure_solve(wv_solver *s, int mode, float tolerance, int max_iterations, int flags, wv_solve_report *out);

#ifdef __cplusplus
}
#endif
#endif
