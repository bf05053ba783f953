// mlp_kernels.h -- launch interface of the value-network kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpuai/mi_mlp.h"
#include "gpuai/mi_playout.h"

namespace gpuai {
namespace mlp {

// C[M][N] (+ epilogue) = A[M][K] B[K][N]; A(m,k) = A[m*sAm + k*sAk], B(k,n) = B[k*sBk + n*sBn]
// with one of each pair of strides equal to 1.  M and N multiples of 64, k_per_split of 32;
// split z covers K range [z*k_per_split, (z+1)*k_per_split) and writes C + z*split_stride.
struct GemmArgs {
  const void *A;
  long sAm, sAk;
  const void *B;
  long sBk, sBn;
  int M, N, k_per_split;
  float *C;               // FP32 output (may be null)
  long split_stride;
  int ldc;                // row stride of C and Cbf
  uint16_t *Cbf;          // BF16 output (may be null)
  const float *bias;      // + bias[n] (may be null)
  int relu;               // max(x, 0)
  const uint16_t *mask;   // keep x where mask[m*ldm + n] > 0 (may be null)
  int ldm;
  int row_valid;          // rows m >= row_valid are written as 0 (0: all rows valid)
};

// Partial slots per 64-row block: b1 at 0-255, b2 256-511, w3 512-767, b3 768, loss 769.
constexpr uint32_t kVecSlots = 1024;

struct AdamArgs {
  float *w, *m, *v, *grads;
  const float *pW1, *pW2;  // split-K partial gradients of W1 ([splits][128*256]) and W2
  uint32_t splits;
  const float *pV;         // column-sum partials ([vsplits][kVecSlots])
  uint32_t vsplits;
  uint16_t *W1t, *W2t, *W2r;
  float lr, beta1, beta2, eps, c1, c2;  // c1 = 1/(1-beta1^t), c2 = 1/(1-beta2^t)
  float omb1, omb2;        // 1 - beta1, 1 - beta2 (formed in double precision on the host)
  int update;              // 0: only sum the gradients and refresh the BF16 copies
  float *loss_out;         // when set, receives the sum of the loss partials (slot 769)
};

hipError_t gemm(const GemmArgs &g, int splits, bool a_bf16, bool b_bf16, hipStream_t stream);
hipError_t encode(const mi_pos *pos, const uint32_t *idx, uint32_t n, uint32_t rows, uint16_t *X,
                  hipStream_t stream);
// Inference in one launch (k_forward): y[r] = network(pos[r]) for r < n, 16 rows per
// workgroup, bit-identical to encode + gemm (both layers) + the output layer.
hipError_t forward(const mi_pos *pos, uint32_t n, const uint16_t *W1t, const float *b1,
                   const uint16_t *W2t, const float *b2, const float *w3, const float *b3, float *y,
                   hipStream_t stream);
// Loss and output-layer gradients over 64-row blocks (rows a multiple of 64); the target of
// row r is targets[tidx ? tidx[r] : r].  Writes dz, dH2 and the partial slots of b2, w3, b3
// and the loss.
hipError_t head_train(const uint16_t *H2, const float *w3, const float *b3, const float *targets,
                      const uint32_t *tidx, uint32_t n, uint32_t rows, float *dz, uint16_t *dH2,
                      float *part, hipStream_t stream);
// Partial column sums of 64-row blocks into part[block * kVecSlots + col].
hipError_t colsum_bf16(const uint16_t *src, uint32_t rows, uint32_t cols, const float *weight,
                       float *part, hipStream_t stream);
hipError_t colsum_f32(const float *src, uint32_t rows, uint32_t cols, const float *weight,
                      float *part, hipStream_t stream);
hipError_t adam(const AdamArgs &a, hipStream_t stream);
hipError_t zero(float *p, uint32_t n, hipStream_t stream);

}  // namespace mlp
}  // namespace gpuai
