/*
 * psst_mlp.h -- C-ABI of the learned warm start: a coordinate MLP
 * (x, y, Mach, ramp angle) -> (rho, u, v, p) that initialises new solves in
 * place of the freestream.  The reference has no counterpart: it always
 * starts from the .par freestream or a .sav restart
 * (psstFiniteVolumeGrid.cpp:641-702, init mode 0/1); the network's output is
 * handed to the solver as such an initial state (psst_set_state), and the
 * solver still runs to its tolerance, so it remains the arbiter of the
 * surrogate's error.
 *
 * Network: n_in inputs, `depth` hidden layers of `width` units with SiLU,
 * n_out linear outputs.  Parameters are kept in FP32 (the optimiser's
 * master copy); the network computes with BF16 weights and BF16 hidden
 * activations and FP32 accumulation on the MFMA units (hidden layers), the
 * input and output layers in FP32 arithmetic on BF16 weights.
 *
 * Parameter vector layout (FP32): for each layer l = 0..depth, W_l stored
 * [fan_in][fan_out] row-major, then b_l [fan_out]; fan_in/fan_out are
 * n_in->width, width->width (depth-1 times), width->n_out.
 */
#ifndef PSST_MLP_H
#define PSST_MLP_H

#include <stdint.h>

#include "psst_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psst_mlp psst_mlp;

/* width: multiple of 64 (<= 1024); depth >= 2 hidden layers; n_in, n_out 1..8.
 * Weights are initialised with a seeded uniform(+-sqrt(6/(fan_in+fan_out)))
 * draw, biases zero. */
psst_status psst_mlp_create(int32_t n_in, int32_t width, int32_t depth, int32_t n_out, uint64_t seed, psst_mlp **out);
int64_t psst_mlp_param_count(const psst_mlp *m);
psst_status psst_mlp_get_params(psst_mlp *m, float *params);
psst_status psst_mlp_set_params(psst_mlp *m, const float *params);
/* Inference: x [n][n_in] -> y [n][n_out] (host arrays). */
psst_status psst_mlp_forward(psst_mlp *m, const float *x, int64_t n, float *y);
/* Mean squared error over the n x n_out outputs and its gradient with
 * respect to every parameter (grad [param_count], FP32). */
psst_status psst_mlp_loss_grad(psst_mlp *m, const float *x, const float *y, int64_t n, double *loss, float *grad);
/* Uploads a training set (x [n][n_in], y [n][n_out]) that stays in HBM. */
psst_status psst_mlp_set_data(psst_mlp *m, const float *x, const float *y, int64_t n);
/* Adam steps on seeded random minibatches of the uploaded set (batch: a
 * multiple of 1024); loss_hist [steps] (may be NULL) receives each step's
 * minibatch loss.  lr follows a cosine decay from lr to lr_end over the
 * steps. */
psst_status psst_mlp_train(psst_mlp *m, int32_t steps, int32_t batch, float lr, float lr_end, uint64_t seed,
                           float *loss_hist);
/* Mean time per training step (ms) of the last psst_mlp_train call and the
 * hidden-layer GEMM FLOPs per step (forward + backward). */
psst_status psst_mlp_train_stats(const psst_mlp *m, double *ms_per_step, double *gemm_flop_per_step);
void psst_mlp_destroy(psst_mlp *m);
const char *psst_mlp_last_error(const psst_mlp *m);

#ifdef __cplusplus
}
#endif
#endif /* PSST_MLP_H */
