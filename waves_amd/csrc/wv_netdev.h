// Solver-side access to the learned warm-start network (wv_net.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct wv_net;
uint16_t* wv_net_input_dev(wv_net* n);        // [cells][8] bf16 input of the first sample
const float* wv_net_qf_dev(wv_net* n);        // fine head output [cells]
const float* wv_net_qc_dev(wv_net* n);        // coarse head output [level-2 cells]
void wv_net_grid(const wv_net* n, int* dims);  // nx, ny, nz, batch
void wv_net_forward_dev(wv_net* n, hipStream_t st);
