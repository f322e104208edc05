// Host-visible launchers of the gfx950 kernels and the multigrid descriptors.
#pragma once
#include "wv_device.h"

#define WV_MAXL 12

// One multigrid level (water/pressure_multigrid.gd:40-69 planning rules).
struct MGLevel {
  int n[3];
  int count;
  int f[3];        // coarsening factor towards the next level
  int aggregate;   // transfers to the next level are piecewise constant
  int width;       // stencil entries of this level's rows
  int off_start;   // first stencil offset (index into MGDev::offsets / 3)
  int64_t vec_off; // first float of this level's 8 vector fields
  int64_t row_off; // first float of this level's k-major rows (levels >= 1)
  int margin[3];   // largest |offset| per axis of the stencil: rows this far from every edge are interior
  int lanes;       // lanes cooperating on one row in the coarse-level sweeps (power of two <= 32)
  int rlanes;      // lanes cooperating on one node in the restriction into this level (power of two <= 32)
  int ga;          // lanes per fine row in the Galerkin stage 1 out of this level (power of two <= 16)
  int box[3];      // extent of one fine row's (A P) box over coarse nodes (Galerkin stage 1)
  int wrap1;       // the stencil reaches at most one period along z (and x when periodic):
                   // a neighbour index is row + delta with one +-period correction
};
struct MGDev {
  int L, perx;
  int tail;                  // first level run by the one-workgroup tail kernel, -1: none
  MGLevel lv[WV_MAXL];
  const int* offsets;        // stencil offsets, 3 ints each
  const signed char* slots;  // per level: 125 delta slots -> stencil index, -1 if absent
  const int* deltas;         // per stencil entry: linear index offset for interior rows
  float* vec;                // node_count * 8 floats: b, x1, x2, active, l1 D^-1, -, r, d
  float* rows;               // coarse operator rows, k-major, FP32
  float* linv;               // last level: dense L^-1 (count^2)
  float* ap;                 // Galerkin scratch: one fine level's (A P) rows, box-major
  uint16_t* rows16;          // BF16 copy of the coarse rows (mixed-precision V-cycle)
  uint16_t* snap16;          // BF16 copy of the 7 level-0 snapshot coefficient fields
  float* snapshot;           // 9 N: preconditioner's fine coefficients, kind, face signature
  float* previous;           // 9 N: previous trial's operator audit copy
  unsigned* state;           // hierarchy audit header + per-trial records
  unsigned* audit_part;      // audit counts per block stripe (AUDIT_STRIPES x 16), summed into state[1..7, 15, 16]
  int16_t* ttab;             // the tail's index tables (wv_mg_tail_tables), null when it has none
  int lag, mode, capacity;
};

int wv_launch_op(const Ctx& s, int op, hipStream_t st);
// pressure-section timing: begin / end stamps around a batch's pressure
// section, and the per-accepted-step record after the acceptance op
void wv_launch_reflect_table(const Ctx& s, hipStream_t st);  // once per solver: the static wall-search table
void wv_launch_press_begin(const Ctx& s, hipStream_t st);
void wv_launch_press_end(const Ctx& s, hipStream_t st);
void wv_launch_step_record(const Ctx& s, hipStream_t st);
// op 8, with the level-0 load of the V-cycle that follows it (and that
// V-cycle's first sweep when wv_mg_first_sweep_fused)
void wv_launch_pcg_update(const Ctx& s, const MGDev& m, hipStream_t st);
bool wv_mg_first_sweep_fused(const MGDev& m);
// batch program: ops 25 + 26 of VOF sweep `sweep` with the finaliser's two
// in-order ledger sums left to op 12's finaliser, and op 12 (+ the sweeps'
// ledger sums when `vof_ledgers`)
void wv_launch_vof_deferred(const Ctx& s, int sweep, hipStream_t st);
void wv_launch_accept(const Ctx& s, bool vof_ledgers, hipStream_t st);
void wv_launch_spmv_dot(const Ctx& s, hipStream_t st);  // ops 5 + 6 + 7 (plain FP32 form)
// Multigrid: build = hierarchy audit + Galerkin rebuild (phase 2);
// cycle = one V-cycle (phase 0 initial / 1 iteration) ending in the fused
// z write-back, partials and PCG finaliser (op 4 or op 9).
void wv_mg_build(const Ctx& s, const MGDev& m, hipStream_t st);
// first level of the one-workgroup tail for this hierarchy on the current device (-1: none)
int wv_mg_tail_level(const MGDev& m);
// 16-bit entries of the tail's index tables (0: the tail has no table form),
// and their one-time fill on the device
int64_t wv_mg_tail_table_shorts(const MGDev& m);
void wv_mg_tail_tables(const MGDev& m, int16_t* tab, hipStream_t st);
void wv_mg_cycle(const Ctx& s, const MGDev& m, int phase, int smoothing, bool load, hipStream_t st);
void wv_launch_pcg_init_partials(const Ctx& s, hipStream_t st);
void wv_launch_spmv_timed(const Ctx& s, hipStream_t st);
// mixed-precision V-cycle: BF16 operators, FP32 vectors and accumulation
void wv_mg_to_bf16(const Ctx& s, const MGDev& m, hipStream_t st);
void wv_mg_cycle_bf16(const Ctx& s, const MGDev& m, int phase, int smoothing, bool load, bool fp64_out, hipStream_t st);
// standalone pressure solve pieces (wv_solve.hip)
void wv_solve_begin(const Ctx& s, bool zero_pressure, hipStream_t st);
void wv_solve_zero_mean(const Ctx& s, hipStream_t st);
void wv_solve_dot_alpha64(const Ctx& s, hipStream_t st);
void wv_solve_true_residual64(const Ctx& s, const MGDev& m, bool replace, int parts, hipStream_t st);
void wv_solve_finish(const Ctx& s, hipStream_t st);
void wv_solve_init_given(const Ctx& s, hipStream_t st);
void wv_solve_volume64(const Ctx& s, int bank, hipStream_t st);
struct wv_net;
void wv_solve_learned_start(const Ctx& s, wv_net* net, hipStream_t st);
