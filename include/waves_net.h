/* Learned warm start for the pressure solve: a 3-D convolutional U-Net with
 * hand-written HIP training and inference kernels (libwaves_hip.so).
 *
 * No counterpart in the reference (which has no learned component).  The
 * network: 4 levels, channels 8-16-32-64, 3x3x3 convolutions computed as
 * implicit GEMMs on bf16 MFMA with fp32 accumulation, bf16 weights and
 * activations, an fp32 master copy of the weights updated by Adam.  Input:
 * 8 channels per cell (F, r/rms r, p/rms p, wet, 4 zeros) as bf16 bits;
 * outputs: a fine field qf (level 0) and a coarse field qc (level 2, prolonged
 * by 4 per axis).  Parameter layout: per layer W[cout][kpad(cin)], row = 27
 * taps x cin (tap = (dz+1)*9 + (dy+1)*3 + (dx+1), channel fastest), then the
 * bias, then zero padding to a multiple of 16.
 */
#ifndef WAVES_NET_H
#define WAVES_NET_H
#include <stdint.h>
#include "waves.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wv_net wv_net;

/* Number of float parameters (the flat layout above). */
int64_t wv_net_param_count(void);
/* A network for grids of nx*ny*nz cells and `batch` samples per pass. */
int wv_net_create(int nx, int ny, int nz, int batch, wv_net **out);
void wv_net_destroy(wv_net *net);
int wv_net_set_params(wv_net *net, const float *params, int64_t count);
int wv_net_get_params(wv_net *net, float *params, int64_t count);
/* Forward pass of `batch` samples: x = batch*cells*8 bf16 bits (host);
 * qf = batch*cells floats, qc = batch*cells(level 2) floats (host). */
int wv_net_forward(wv_net *net, const uint16_t *x, float *qf, float *qc);
/* Loss (mean over samples of sum_wet (qf + up4 qc - t)^2 / sum_wet t^2, wet =
 * input channel 3 > 0.5) and its gradient with respect to every parameter. */
int wv_net_loss_grad(wv_net *net, const uint16_t *x, const float *target, float *loss, float *grad);
/* Upload `count` training samples (x as above, target = count*cells floats). */
int wv_net_set_dataset(wv_net *net, const uint16_t *x, const float *target, int count);
/* `steps` Adam steps (lr, beta1 0.9, beta2 0.999, eps 1e-8) on batches drawn
 * from dataset samples [0, train_count) with a seeded generator; per-step
 * losses are written to `losses` (steps floats, may be NULL). */
int wv_net_train(wv_net *net, int steps, int train_count, float lr, uint64_t seed, float *losses);
/* Mean loss over dataset samples [first, first + count) (batches of `batch`). */
int wv_net_eval(wv_net *net, int first, int count, float *loss);
const char *wv_net_last_error(wv_net *net);
/* The network's layers in forward order: count, and layer l's grid level
 * (0 = full grid, each level halves every axis rounding up), input channels
 * at that level (ca), input channels upsampled from the next level (cb) and
 * output channels. */
int wv_net_layer_count(void);
int wv_net_layer_info(int l, int *level, int *ca, int *cb, int *cout);
/* One forward pass of the batch already in the network's input buffer (the
 * last wv_net_forward's), then every layer's convolution launched `reps` times
 * on the network's stream and timed with HIP events: us[l] = mean device
 * microseconds per launch of layer l (count = wv_net_layer_count()). */
int wv_net_time_layers(wv_net *net, int reps, float *us, int count);

/* Attach a network (created with the solver's grid and batch 1) for the
 * learned warm start of wv_pressure_solve mode 2; NULL detaches.  The solver
 * does not own the network. */
int wv_solver_attach_net(wv_solver *s, wv_net *net);

#ifdef __cplusplus
}
#endif
#endif
