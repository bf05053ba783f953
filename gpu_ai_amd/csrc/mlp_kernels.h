// mlp_kernels.h -- launch interface of the value-network kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpuai/mi_mlp.h"
#include "gpuai/mi_playout.h"

namespace gpuai {
namespace mlp {

// Rows per workgroup of k_train_rows (a multiple of 16, dividing 64): the training batch is a
// multiple of 64 rows.
constexpr uint32_t kTrainRows = 32;

// Rows per split of the weight-gradient products (split-K over the batch; a multiple of 32).
constexpr uint32_t kWSplitRows = 128;

// Partial slots per row block of k_train_rows: b1 at 0-255, b2 256-511, w3 512-767, b3 768,
// loss 769.
constexpr uint32_t kVecSlots = 1024;

// One training step's row pass (k_train_rows), kTrainRows rows per workgroup: row r < n is
// pos[idx ? idx[r] : r] with target targets[idx ? idx[r] : r]; rows n and up (to the multiple
// of kTrainRows) are padding.  Writes the transposed activations X^T, H1^T, dH2^T, dH1^T ([feature][ld],
// the step's rows at 0..rows-1) and block b's vector partials at pV + b * kVecSlots.
struct TrainArgs {
  const mi_pos *pos;
  const uint32_t *idx;
  const float *targets;
  uint32_t n;
  uint32_t ld;  // row stride of the transposed activations (>= rows, a multiple of 64)
  const uint16_t *W1t, *W2t, *W2r;
  const float *b1, *b2, *w3, *b3;
  uint16_t *XT, *H1T, *dH2T, *dH1T;
  float *pV;
};

// Weight gradients over the step's rows (k_wgrad_part: split-K partials in P, `splits` of them,
// one per 256 rows), then Adam with every gradient summed from its partials (k_adam).  rows = 0
// takes every gradient as 0.
struct AdamArgs {
  float *w, *m, *v, *grads;
  const uint16_t *XT, *H1T, *dH2T, *dH1T;  // as the row pass wrote them
  uint32_t rows, ld;
  float *P;          // weight-gradient partials: [splits][W1 32768 | W2 65536]
  uint32_t splits;   // ceil(rows / 256)
  const float *pV;   // the row pass's vector partials, `blocks` of them (rows / kTrainRows)
  uint32_t blocks;
  uint16_t *W1t, *W2t, *W2r;
  float lr, beta1, beta2, eps, c1, c2;  // c1 = 1/(1-beta1^t), c2 = 1/(1-beta2^t)
  float omb1, omb2;        // 1 - beta1, 1 - beta2 (formed in double precision on the host)
  int update;              // 0: only sum the gradients and refresh the BF16 copies
  float *loss_out;         // when set, receives the sum of the loss partials (slot 769)
};

// Inference in one launch (k_forward): y[r] = network(pos[r]) for r < n, 16 rows per workgroup.
hipError_t forward(const mi_pos *pos, uint32_t n, const uint16_t *W1t, const float *b1,
                   const uint16_t *W2t, const float *b2, const float *w3, const float *b3, float *y,
                   hipStream_t stream);
hipError_t train_rows(const TrainArgs &a, uint32_t rows, hipStream_t stream);
hipError_t wgrad_adam(const AdamArgs &a, hipStream_t stream);
hipError_t zero(float *p, uint32_t n, hipStream_t stream);

}  // namespace mlp
}  // namespace gpuai
