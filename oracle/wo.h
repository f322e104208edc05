/*
 * waves oracle -- CPU restatement of the reference solver's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or as the
 * timed CPU baseline). The product (waves_amd/, libwaves_hip.so) never links,
 * loads or calls it.
 *
 * The restatement follows the GLSL kernels of the reference op by op, one
 * invocation at a time, with the 128-lane workgroup tree reductions replayed
 * in the same order, so that single-op outputs are reproducible and the
 * summation order of every reduction is the reference's:
 *   water/shaders/compute/pressure_core.glslinc, plic_math.glslinc,
 *   pressure_geometry.glslinc, height_normal.glslinc, pressure_pass.glslinc,
 *   pressure_iterations.glslinc, coupled_motion.glslinc, coastal_target.glslinc,
 *   coastal_pass.glslinc, pressure_multigrid.glslinc, pressure_mg_coarse.glslinc,
 *   pressure_mg_cache_arguments.glsl, pressure_hierarchy_state.glsl
 * and the host sequencing of water/pressure_gpu.gd, water/coupled_gpu.gd,
 * water/coastal_gpu.gd, water/pressure_multigrid.gd,
 * water/pressure_hierarchy_state.gd.
 *
 * Built with -ffp-contract=off: the reference marks its exact-cancellation
 * expressions `precise` (coupled_motion.glslinc:47-48, 393-396).
 */
#ifndef WAVES_ORACLE_H
#define WAVES_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct WO WO;

/* pc: the 32 push-constant floats of water/pressure_gpu.gd:68,
 * water/coupled_gpu.gd:18-21, water/coastal_gpu.gd:23-24 (index 3 = op and
 * index 15 = quota are overwritten per dispatch).
 * opts: [0] right_absorption (coastal_gpu.gd:12), [1] mg smoothing (3),
 *       [2] mg lag (8), [3] mg aggregate_from (2), [4] mg maintenance mode
 *       (0 rebuild, 1 structure, 2 lagged), [5] PCG iterations per batch (8),
 *       [6] alongshore perturbation amplitude [m], [7] perturbation seed. */
WO *wo_create(const float *pc, const float *opts);
void wo_destroy(WO *s);
/* field ids: see WV_FIELD_* in include/waves.h (same numbering). */
int64_t wo_field_count(WO *s, int field);
int wo_get(WO *s, int field, void *dst, int64_t count);
int wo_set(WO *s, int field, const void *src, int64_t count);
/* one dispatch of `op` over `groups` 128-lane groups (reference numbering). */
void wo_op(WO *s, int op, int64_t groups);
void wo_mg_record(WO *s, int build, int initial);
void wo_set_quota(WO *s, float quota);
/* order of the Galerkin scatter's additions (MG op 3): 0 ascending entry
 * index, 1 descending.  The reference adds with float compare-and-swap
 * atomics in no fixed order (pressure_multigrid.glslinc:128-134), so both are
 * executions of the reference; the switch measures what the order alone does
 * to a trajectory. */
void wo_set_scatter_order(WO *s, int reverse);
/* initial command list: ops 0,40,(perturb),17,18,16,19,41,42,49,85,86 */
void wo_initial(WO *s);
/* one full batch of water/coupled_gpu.gd:38-56 + coastal ops, quota as in
 * pressure_pass.glslinc:20 (>=0 accepted steps, <0 -deadline). */
void wo_batch(WO *s, float quota);
int wo_levels(WO *s);
/* level descriptors: 16 ints per level as the reference's `meta` */
int wo_level_meta(WO *s, int level, int *dst16);
/* apply the V-cycle preconditioner to r (N floats) -> z, with the current
 * hierarchy, outside of any control gating */
void wo_precondition(WO *s, const float *r, float *z);
/* closed forms exposed for the tests */
float wo_cube_volume(const float *normal, float alpha);
float wo_inverse_volume(const float *normal, float fraction);
#ifdef __cplusplus
}
#endif
#endif
