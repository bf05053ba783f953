/*
 * mi_mlp.h -- C ABI of the learned leaf value: a 128-256-256-1 MLP (ReLU, ReLU, tanh) over the
 * position seen by the side to move, trained and evaluated by hand-written gfx950 kernels
 * (BF16 operands on MFMA v_mfma_f32_16x16x32_bf16, FP32 accumulation, FP32 master weights and
 * Adam moments).  No counterpart in the reference, whose leaves are only ever played out
 * (src/mcts.cpp:57-61); plain playouts (mi_playout.h) stay the fallback and the yardstick.
 *
 * Input encoding (128 values, 0 or 1): four planes of 32 squares in the frame of the side to
 * move (its men, its kings, the opponent's men, the opponent's kings); for player 2 the board is
 * rotated by 180 degrees (square s -> 31 - s), so both sides see themselves moving "up".
 * Output: the expected result for the side to move, in [-1, 1] (win 1, draw 0, loss -1).
 *
 * Parameters, FP32, in this order (99,073 values):
 *   W1[128][256], b1[256], W2[256][256], b2[256], w3[256], b3[1]
 * (W[i][o]: input i to output o).  The forward pass rounds inputs, weights and stored
 * activations to BF16 (round to nearest even) and accumulates in FP32.
 */
#ifndef GPUAI_MI_MLP_H
#define GPUAI_MI_MLP_H

#include <stdint.h>

#include "gpuai/mi_playout.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MI_MLP_INPUTS 128
#define MI_MLP_HIDDEN 256
#define MI_MLP_PARAMS 99073
#define MI_MLP_OFF_W1 0
#define MI_MLP_OFF_B1 32768
#define MI_MLP_OFF_W2 33024
#define MI_MLP_OFF_B2 98560
#define MI_MLP_OFF_W3 98816
#define MI_MLP_OFF_B3 99072

typedef struct mi_mlp mi_mlp;

/* Creates a network on the context's device for batches of up to max_batch positions, with
 * uniform He initialisation drawn from Philox key (seed, tensor), counter (i >> 2, 0, 3, 0)
 * (biases 0), Adam beta1 0.9, beta2 0.999, eps 1e-8. */
int mi_mlp_create(mi_ctx *ctx, uint32_t max_batch, uint32_t seed, mi_mlp **mlp);
int mi_mlp_destroy(mi_mlp *mlp);

/* FP32 master parameters (MI_MLP_PARAMS values).  set also clears the Adam state. */
int mi_mlp_get_params(mi_mlp *mlp, float *params);
int mi_mlp_set_params(mi_mlp *mlp, const float *params);

/* Values of n positions for the side to move (n <= max_batch). */
int mi_mlp_forward(mi_mlp *mlp, const mi_pos *pos, uint32_t n, float *values);

/* Gradient of the mean squared error (1/n) sum (value - target)^2 with respect to every
 * parameter (MI_MLP_PARAMS values, the parameter layout), without updating anything. */
int mi_mlp_gradients(mi_mlp *mlp, const mi_pos *pos, const float *targets, uint32_t n,
                     float *grads, float *loss);

/* One Adam step (learning rate lr) on one batch; `loss` (may be NULL) receives the batch MSE
 * before the step. */
int mi_mlp_train_step(mi_mlp *mlp, const mi_pos *pos, const float *targets, uint32_t n,
                      float lr, float *loss);

/* Adam moments (MI_MLP_PARAMS each, may be NULL) and the step count. */
int mi_mlp_adam_state(mi_mlp *mlp, float *m, float *v, uint32_t *step);

/* Device-resident training set (positions and targets copied once), and one epoch over it in
 * batches of `batch` (a multiple of 64, <= max_batch) in an order drawn from Philox key
 * (shuffle_seed, epoch), counter (i >> 2, 0, 3, 1); a trailing partial batch is dropped.
 * mean_loss receives the mean of the batch losses; ms the device time of the epoch. */
int mi_mlp_set_dataset(mi_mlp *mlp, const mi_pos *pos, const float *targets, uint32_t n);
int mi_mlp_train_epoch(mi_mlp *mlp, uint32_t batch, float lr, uint32_t shuffle_seed,
                       uint32_t epoch, float *mean_loss, float *ms);

/* Checkpoint: parameters, Adam moments and step, in a little-endian binary file. */
int mi_mlp_save(mi_mlp *mlp, const char *path);
int mi_mlp_load(mi_mlp *mlp, const char *path);

#ifdef __cplusplus
}
#endif
#endif
