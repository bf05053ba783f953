// mlp_kernels.hip -- gfx950 kernels of the 128-256-256-1 value network.
//
// MFMA v_mfma_f32_16x16x32_bf16 throughout: lane l supplies row l%16 of A and column l%16 of B
// for eight K values of its lane group l/16 (the same eight for A and B, which is all the dot
// product needs: every operand is stored with K contiguous, so each is one 16-byte read), and
// receives D[4*(l/16)+v][l%16], v = 0..3.
//
// A training step is three launches.  k_train_rows runs the forward and backward pass of
// kTrainRows (32) rows per workgroup up to the activation gradients, with the activations in
// LDS, and writes what the weight gradients need transposed (X^T, H1^T, dH2^T, dH1^T:
// [feature][row], rows as K) plus each block's partial sums of the vector gradients.  Every
// workgroup reads all of the BF16 weights (320 KB) from L2; 32 rows per workgroup measured
// 36.2 us per 4096-row step against 38.6 with 64 rows and 39.3 with 16.  k_wgrad_part forms
// 64x64 tiles of dW1 and dW2 over kWSplitRows-row splits of the batch (split-K: 24 tiles x 32
// splits of 128 rows for 4096 rows; 128 measured 31.6 us per step, 256 32.1, 512 32.3), and
// k_adam sums every gradient's partials in a fixed order and applies Adam, refreshing the BF16
// copies.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_kernels.h"

namespace gpuai {
namespace mlp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Workgroup barrier for exchanges through LDS only: this wave's LDS accesses complete, then
// s_barrier.  __syncthreads() also waits for the wave's global stores to complete, which costs a
// memory round trip after every store of activations or partial sums; nothing here reads back
// what the workgroup stored to global memory.  The memory clobber keeps the compiler from moving
// memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// The four input planes of a position in the mover's frame (own men, own kings, opponent's men,
// opponent's kings), 32 squares each.
__device__ __forceinline__ uint32_t plane_bits(const mi_pos &p, int plane) {
  uint32_t own = p.turn ? __builtin_bitreverse32(p.p2) : p.p1;
  uint32_t opp = p.turn ? __builtin_bitreverse32(p.p1) : p.p2;
  uint32_t kings = p.turn ? __builtin_bitreverse32(p.kings) : p.kings;
  kings &= own | opp;
  return plane == 0 ? (own & ~kings) : plane == 1 ? (own & kings)
       : plane == 2 ? (opp & ~kings) : (opp & kings);
}

// Inference in one launch: encoding, both hidden layers and the output layer for FR rows per
// workgroup, with the activations in LDS: H1 = bf(relu(X W1 + b1)), H2 = bf(relu(H1 W2 + b2)),
// y = tanh(H2 . w3 + b3), the output layer summing four columns per lane and then over the wave
// with a butterfly.  Wave w computes output columns 64w..64w+63 of each layer as four 16x16
// tiles; the B operands (W1t, W2t: [out][in], in contiguous) are read from global memory (they
// stay in L2), one 16-byte load per lane per tile and slice.
constexpr int FR = 16, XP = 128 + 8, HP = 256 + 8;  // rows; LDS row pitches (BF16 elements)

__global__ void __launch_bounds__(256) k_forward(const mi_pos *__restrict__ pos, uint32_t n,
                                                const uint16_t *__restrict__ W1t,
                                                const float *__restrict__ b1,
                                                const uint16_t *__restrict__ W2t,
                                                const float *__restrict__ b2,
                                                const float *__restrict__ w3, const float *b3,
                                                float *__restrict__ y) {
  __shared__ __attribute__((aligned(16))) uint16_t X[FR][XP];
  __shared__ __attribute__((aligned(16))) uint16_t H[FR][HP];  // H1, then H2
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t r0 = blockIdx.x * FR;
  {  // encode: thread t -> row t/16, plane (t/4)%4, squares 8(t%4)..+7
    const int row = t >> 4, plane = (t >> 2) & 3, s0 = (t & 3) * 8;
    uint32_t bits = 0;
    if (r0 + row < n) bits = plane_bits(pos[r0 + row], plane);
    bits >>= s0;
    uint32_t w[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
      w[j] = ((bits >> (2 * j)) & 1u ? 0x3F80u : 0u) | ((bits >> (2 * j + 1)) & 1u ? 0x3F800000u : 0u);
    *(uint4 *)&X[row][plane * 32 + s0] = uint4{w[0], w[1], w[2], w[3]};
  }
  __syncthreads();
  const int kg = (lane >> 4) * 8;  // this lane's eight K values within a slice
  f32x4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < 128; k0 += 32) {  // layer 1: X[16][128] . W1
    const bf16x8 a = *(const bf16x8 *)&X[lane & 15][k0 + kg];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int col = wave * 64 + c * 16 + (lane & 15);
      const bf16x8 b = *(const bf16x8 *)(W1t + col * 128 + k0 + kg);
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[c], 0, 0, 0);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int col = wave * 64 + c * 16 + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      float x = acc[c][v] + b1[col];
      x = x > 0.f ? x : 0.f;
      H[4 * (lane >> 4) + v][col] = to_bf(x);
      acc[c][v] = 0.f;
    }
  }
  __syncthreads();
  for (int k0 = 0; k0 < 256; k0 += 32) {  // layer 2: H1[16][256] . W2
    const bf16x8 a = *(const bf16x8 *)&H[lane & 15][k0 + kg];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int col = wave * 64 + c * 16 + (lane & 15);
      const bf16x8 b = *(const bf16x8 *)(W2t + col * 256 + k0 + kg);
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[c], 0, 0, 0);
    }
  }
  __syncthreads();  // every wave has read H1
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int col = wave * 64 + c * 16 + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      float x = acc[c][v] + b2[col];
      x = x > 0.f ? x : 0.f;
      H[4 * (lane >> 4) + v][col] = to_bf(x);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < FR / 4; i++) {  // output layer: wave w, rows 4w..4w+3
    const int row = wave * (FR / 4) + i;
    float part = 0.f;
#pragma unroll
    for (int c = 0; c < 4; c++) part += from_bf(H[row][lane * 4 + c]) * w3[lane * 4 + c];
    for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off);
    if (lane == 0 && r0 + row < n) y[r0 + row] = tanhf(part + b3[0]);
  }
}

// ---------------------------------------------------------------------------------------------
// Training.  Row r of the batch is pos[idx ? idx[r] : r] with target targets[idx ? idx[r] : r];
// rows n and up (to the multiple of kTrainRows) are padding: zero inputs and no loss.  The
// arithmetic: H1 = bf(relu(X W1 + b1)), H2 = bf(relu(H1 W2 + b2)), y = tanh(H2 . w3 + b3),
// dz = 2 (y - t)(1 - y^2) / n, dH2 = bf(dz w3) where H2 > 0, dH1 = bf(dH2 W2^T) where H1 > 0;
// dW1 = X^T dH1, dW2 = H1^T dH2, b1 and b2 the column sums of dH1 and dH2, w3 = sum of dz H2,
// b3 = sum of dz, loss = sum of (y - t)^2 / n.

constexpr int KR = (int)kTrainRows;  // rows per workgroup of k_train_rows
constexpr int RT = KR / 16;             // row tiles
constexpr int TP = KR + 8;  // pitch of the transposed staging tile (XP and HP as k_forward's)
static_assert(KR % 16 == 0 && 64 % KR == 0, "rows per workgroup: 16, 32 or 64");

// Copies the transposed staging tile T ([f][row], f < features) to G ([f][rows_total], this
// block's KR rows at r0): 2 KR bytes per feature in 16-byte stores.
__device__ __forceinline__ void store_transposed(const uint16_t (*T)[TP], int features,
                                                 uint16_t *__restrict__ G, uint32_t ld,
                                                 uint32_t r0) {
  constexpr int U = KR / 8;  // 16-byte units per feature
  for (int i = threadIdx.x; i < features * U; i += 512) {
    const int f = i / U, u = i % U;
    *(uint4 *)(G + (size_t)f * ld + r0 + 8 * u) = *(const uint4 *)&T[f][8 * u];
  }
}

// The forward and backward pass of KR rows: wave w (of eight) owns columns 32w..32w+31 of each
// 256-wide layer (two 16x16 column tiles, RT row tiles).  Each workgroup reads the whole of the
// BF16 weights (320 KB), so the rows per workgroup set the weight traffic through L2.  A operands
// come from LDS ([row][k]), B operands from the BF16 weight copies in global memory (W1t, W2t:
// [out][in]; W2r: [in][out]).
// Every global load is issued at the start (a dependent load waits out the memory latency, and
// a layer's panel loaded slice by slice would wait it out once per 32-deep slice); the pass then
// only computes and stores.
__global__ void __launch_bounds__(512) k_train_rows(TrainArgs a) {
  __shared__ __attribute__((aligned(16))) uint16_t Xs[KR][XP];
  __shared__ __attribute__((aligned(16))) uint16_t H1s[KR][HP];
  __shared__ __attribute__((aligned(16))) uint16_t dH2s[KR][HP];
  __shared__ __attribute__((aligned(16))) uint16_t T[256][TP];  // transposed staging
  __shared__ float red[8][KR];
  __shared__ float dzs[KR];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int lr = lane & 15, kg = (lane >> 4) * 8;
  const uint32_t r0 = blockIdx.x * KR, ld = a.ld;
  float *pv = a.pV + (size_t)blockIdx.x * kVecSlots;
  // the wave's columns, rotated by the block index: all workgroups read the same weights at the
  // same time, and the rotation spreads their requests over different lines of L2
  const int cw = ((wave + (int)blockIdx.x) & 7) * 32;
  // Every global load the pass needs, issued before any is used and unconditionally (padding rows
  // load row n - 1 and ignore it), in an order that lets each wait cover only what it needs --
  // loads complete in issue order as far as the wait counter goes: the wave's W1t panel with the
  // rows' dataset indices, then the positions and targets they point to, then the W2t and W2r
  // panels and the biases.  The empty asm statements keep the compiler from reordering the
  // groups (it would hoist the index-dependent address arithmetic, and its wait, to the top).
  constexpr int TPR = 512 / KR, TPP = TPR / 4, SPT = 32 / TPP;  // threads per row and plane,
  const int erow = t / TPR;                                       // squares per thread
  const bool evalid = r0 + erow < a.n;
  const uint32_t nlast = a.n - 1u;
  bf16x8 w1p[4][2], w2p[8][2], w2rp[8][2];  // this wave's weight panels: [slice][column tile]
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int c = 0; c < 2; c++)
      w1p[s][c] = *(const bf16x8 *)(a.W1t + (cw + c * 16 + lr) * 128 + s * 32 + kg);
  __asm__ volatile("" ::: "memory");
  uint32_t esrc = min(r0 + erow, nlast), tsrc = min(r0 + lane, nlast);
  if (a.idx) {
    esrc = a.idx[esrc];
    tsrc = a.idx[tsrc];
  }
  __asm__ volatile("" ::: "memory");
  const mi_pos epos = a.pos[esrc];
  const float tgt = a.targets[tsrc];
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int s = 0; s < 8; s++)
#pragma unroll
    for (int c = 0; c < 2; c++) {
      w2p[s][c] = *(const bf16x8 *)(a.W2t + (cw + c * 16 + lr) * 256 + s * 32 + kg);
      w2rp[s][c] = *(const bf16x8 *)(a.W2r + (cw + c * 16 + lr) * 256 + s * 32 + kg);
    }
  float b1c[2], b2c[2], w3c[2];
#pragma unroll
  for (int c = 0; c < 2; c++) {
    b1c[c] = a.b1[cw + c * 16 + lr];
    b2c[c] = a.b2[cw + c * 16 + lr];
    w3c[c] = a.w3[cw + c * 16 + lr];
  }
  const float b3 = a.b3[0];
  {  // encode: thread t -> row t/TPR, plane (t/TPP)%4, squares SPT(t%TPP)..+SPT-1
    const int row = erow, plane = (t / TPP) & 3, s0 = (t % TPP) * SPT;
    uint32_t bits = 0;
    if (evalid) bits = plane_bits(epos, plane) >> s0;
    uint32_t w[SPT / 2];
#pragma unroll
    for (int j = 0; j < SPT / 2; j++)
      w[j] = ((bits >> (2 * j)) & 1u ? 0x3F80u : 0u) | ((bits >> (2 * j + 1)) & 1u ? 0x3F800000u : 0u);
#pragma unroll
    for (int j = 0; j < SPT / 2; j += 4)
      *(uint4 *)&Xs[row][plane * 32 + s0 + 2 * j] = uint4{w[j], w[j + 1], w[j + 2], w[j + 3]};
#pragma unroll
    for (int s = 0; s < SPT; s++) T[plane * 32 + s0 + s][row] = ((bits >> s) & 1u) ? 0x3F80u : 0u;
  }
  lds_barrier();
  store_transposed(T, 128, a.XT, ld, r0);
  f32x4 acc[RT][2];
  // ---- layer 1: X[KR][128] . W1 -> H1 (row-major and transposed)
#pragma unroll
  for (int r = 0; r < RT; r++)
#pragma unroll
    for (int c = 0; c < 2; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 4; s++) {
#pragma unroll
    for (int r = 0; r < RT; r++) {
      const bf16x8 x = *(const bf16x8 *)&Xs[r * 16 + lr][s * 32 + kg];
#pragma unroll
      for (int c = 0; c < 2; c++)
        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, w1p[s][c], acc[r][c], 0, 0, 0);
    }
  }
  lds_barrier();  // X^T has left the staging tile
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int col = cw + c * 16 + lr;
    const float bias = b1c[c];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = r * 16 + 4 * (lane >> 4) + v;
        const float x = acc[r][c][v] + bias;
        const uint16_t h = to_bf(x > 0.f ? x : 0.f);
        H1s[row][col] = h;
        T[col][row] = h;
      }
  }
  lds_barrier();
  store_transposed(T, 256, a.H1T, ld, r0);
  // ---- layer 2: H1[KR][256] . W2 -> H2 (registers), the output layer, dz
#pragma unroll
  for (int r = 0; r < RT; r++)
#pragma unroll
    for (int c = 0; c < 2; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 8; s++) {
#pragma unroll
    for (int r = 0; r < RT; r++) {
      const bf16x8 h = *(const bf16x8 *)&H1s[r * 16 + lr][s * 32 + kg];
#pragma unroll
      for (int c = 0; c < 2; c++)
        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h, w2p[s][c], acc[r][c], 0, 0, 0);
    }
  }
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const float bias = b2c[c];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const float x = acc[r][c][v] + bias;
        acc[r][c][v] = from_bf(to_bf(x > 0.f ? x : 0.f));  // H2
      }
  }
#pragma unroll
  for (int r = 0; r < RT; r++)
#pragma unroll
    for (int v = 0; v < 4; v++) {  // this wave's share of the row's dot product with w3
      float d = 0.f;
#pragma unroll
      for (int c = 0; c < 2; c++) d += acc[r][c][v] * w3c[c];
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) d += __shfl_xor(d, off);
      if (lr == 0) red[wave][r * 16 + 4 * (lane >> 4) + v] = d;
    }
  lds_barrier();
  if (wave == 0) {  // output, loss and dz per row (lane = row); the b3 and loss partials
    const uint32_t row = r0 + lane;
    float d = 0.f, l = 0.f;
    {
      float dot = 0.f;
#pragma unroll
      for (int w = 0; w < 8; w++) dot += red[w][lane % KR];
      const float y = tanhf(dot + b3);
      if (lane < KR && row < a.n) {
        const float e = y - tgt;
        d = 2.f * e * (1.f - y * y) / (float)a.n;
        l = e * e / (float)a.n;
      }
      if (lane < KR) dzs[lane] = d;
    }
    for (int off = 32; off >= 1; off >>= 1) {
      d += __shfl_xor(d, off);
      l += __shfl_xor(l, off);
    }
    if (lane == 0) {
      pv[768] = d;
      pv[769] = l;
    }
  }
  lds_barrier();
  {  // dH2 = bf(dz w3) where H2 > 0 (row-major and transposed); b2 and w3 partials
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int col = cw + c * 16 + lr;
      float sb2 = 0.f, sw3 = 0.f;
#pragma unroll
      for (int r = 0; r < RT; r++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int row = r * 16 + 4 * (lane >> 4) + v;
          const float d = dzs[row], h = acc[r][c][v];
          const uint16_t g = to_bf(h > 0.f ? d * w3c[c] : 0.f);
          dH2s[row][col] = g;
          T[col][row] = g;
          sb2 += from_bf(g);
          sw3 += d * h;
        }
      sb2 += __shfl_xor(sb2, 16);
      sw3 += __shfl_xor(sw3, 16);
      sb2 += __shfl_xor(sb2, 32);
      sw3 += __shfl_xor(sw3, 32);
      if (lane < 16) {
        pv[256 + col] = sb2;
        pv[512 + col] = sw3;
      }
    }
  }
  lds_barrier();
  store_transposed(T, 256, a.dH2T, ld, r0);
  // ---- dH1 = bf(dH2 . W2^T) where H1 > 0 (transposed); b1 partials
#pragma unroll
  for (int r = 0; r < RT; r++)
#pragma unroll
    for (int c = 0; c < 2; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 8; s++) {
#pragma unroll
    for (int r = 0; r < RT; r++) {
      const bf16x8 g = *(const bf16x8 *)&dH2s[r * 16 + lr][s * 32 + kg];
#pragma unroll
      for (int c = 0; c < 2; c++)
        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(g, w2rp[s][c], acc[r][c], 0, 0, 0);
    }
  }
  lds_barrier();  // dH2^T has left the staging tile
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const int col = cw + c * 16 + lr;
    float sb1 = 0.f;
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = r * 16 + 4 * (lane >> 4) + v;
        const uint16_t g = to_bf(H1s[row][col] ? acc[r][c][v] : 0.f);  // H1 >= 0: nonzero is > 0
        T[col][row] = g;
        sb1 += from_bf(g);
      }
    sb1 += __shfl_xor(sb1, 16);
    sb1 += __shfl_xor(sb1, 32);
    if (lane < 16) pv[col] = sb1;
  }
  lds_barrier();
  store_transposed(T, 256, a.dH1T, ld, r0);
}

// One Adam step of parameter p with gradient g, from its values w, m, v (loaded by the caller
// beforehand, together with the gradient's partials); returns the new value.
__device__ __forceinline__ float adam_step(const AdamArgs &a, uint32_t p, float g, float w, float m,
                                           float v) {
  a.grads[p] = g;
  if (a.update) {
    m = a.beta1 * m + a.omb1 * g;
    v = a.beta2 * v + a.omb2 * g * g;
    a.m[p] = m;
    a.v[p] = v;
    w -= a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
    a.w[p] = w;
  }
  return w;
}

// The sum of n partials at p, p + stride, ... in order, their loads issued 16 at a time.
__device__ __forceinline__ float sum_partials(const float *p, size_t stride, uint32_t n) {
  float g = 0.f;
  for (uint32_t s0 = 0; s0 < n; s0 += 16u) {
    float x[16];
#pragma unroll
    for (uint32_t j = 0; j < 16u; j++) x[j] = s0 + j < n ? p[(s0 + j) * stride] : 0.f;
#pragma unroll
    for (uint32_t j = 0; j < 16u; j++) g += x[j];
  }
  return g;
}

// Weight-gradient partials: workgroup (tile, split) forms one 64x64 tile of dW2 (tiles 0..15,
// [256 in][256 out]) or dW1 (16..23, [128][256]) over the split's rows [S split, +S) of the
// batch, dW1[j][i] = sum_r X^T[j][r] dH1^T[i][r], dW2[i][o] = sum_r H1^T[i][r] dH2^T[o][r], and
// writes it to P + split * kWParams (W1 at 0, W2 at 32768, each [in][out]); S = kWSplitRows.  The
// split's operand rows (64 features x S rows each) come in with one round of 16-byte loads (S/16
// per thread in flight) into LDS; rows past the batch read as zero.  Wave w computes the 32x32
// quadrant (w / 2, w % 2) as 2 x 2 MFMA tiles over S/32 slices of 32.
constexpr uint32_t kWSplit = kWSplitRows, kWTiles = 24, kWParams = 32768 + 65536;
constexpr int WP = (int)kWSplit + 8;  // LDS pitch of the staged operand rows
constexpr uint32_t kWLoads = kWSplit / 32;  // 16-byte loads per thread and operand
static_assert(kWSplit % 32 == 0, "split rows: a multiple of 32");

__global__ void __launch_bounds__(256) k_wgrad_part(AdamArgs a) {
  __shared__ __attribute__((aligned(16))) uint16_t As[64][WP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64][WP];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t tile = blockIdx.x, split = blockIdx.y;
  const bool w1 = tile >= 16u;
  const uint32_t tb = w1 ? tile - 16u : tile;
  const uint32_t m0 = (tb >> 2) * 64u, n0 = (tb & 3u) * 64u;  // output rows (in), columns (out)
  const uint16_t *A = w1 ? a.XT : a.H1T;
  const uint16_t *B = w1 ? a.dH1T : a.dH2T;
  const uint32_t k0 = split * kWSplit;
  constexpr uint32_t UR = kWSplit / 8;  // 16-byte units per operand row
  uint4 va[kWLoads], vb[kWLoads];
#pragma unroll
  for (uint32_t j = 0; j < kWLoads; j++) {
    const uint32_t u = t + 256u * j, row = u / UR, k = k0 + (u % UR) * 8u;
    const bool in = k < a.rows;
    va[j] = in ? *(const uint4 *)(A + (size_t)(m0 + row) * a.ld + k) : uint4{0u, 0u, 0u, 0u};
    vb[j] = in ? *(const uint4 *)(B + (size_t)(n0 + row) * a.ld + k) : uint4{0u, 0u, 0u, 0u};
  }
#pragma unroll
  for (uint32_t j = 0; j < kWLoads; j++) {
    const uint32_t u = t + 256u * j, row = u / UR, c = (u % UR) * 8u;
    *(uint4 *)&As[row][c] = va[j];
    *(uint4 *)&Bs[row][c] = vb[j];
  }
  __syncthreads();
  const uint32_t qr = (wave >> 1) * 32u, qc = (wave & 1u) * 32u;
  const uint32_t lr = lane & 15u, kg = (lane >> 4) * 8u;
  f32x4 acc[2][2];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int c = 0; c < 2; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (uint32_t s = 0; s < kWSplit / 32; s++) {
    bf16x8 x[2], y[2];
#pragma unroll
    for (int r = 0; r < 2; r++) x[r] = *(const bf16x8 *)&As[qr + r * 16 + lr][s * 32 + kg];
#pragma unroll
    for (int c = 0; c < 2; c++) y[c] = *(const bf16x8 *)&Bs[qc + c * 16 + lr][s * 32 + kg];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
      for (int c = 0; c < 2; c++)
        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x[r], y[c], acc[r][c], 0, 0, 0);
  }
  float *P = a.P + (size_t)split * kWParams + (w1 ? 0u : 32768u);
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        P[(m0 + qr + r * 16 + 4 * (lane >> 4) + v) * 256u + n0 + qc + c * 16 + lr] = acc[r][c][v];
}

// Adam over every parameter.  Each gradient is the sum of its partials in a fixed order: W1 and
// W2 over the weight-gradient splits (P), in split order; b1, b2, w3 and b3 over the row pass's
// row blocks (pV; slot 769 holds the loss), as eight sums over consecutive eighths of the blocks,
// added in order (so each thread's loads are one round: 128 blocks at 4096 rows).  Workgroups
// take 16x16 tiles of W1 ([128][256]) and W2 ([256][256]), then the vector parameters, 32 slots
// per workgroup; the BF16 copies are refreshed, the transposed ones (W1t, W2t: [out][in])
// through an LDS tile so that their stores are contiguous.
constexpr uint32_t AT = 16, kAdamW1 = (128 / AT) * (256 / AT), kAdamW2 = (256 / AT) * (256 / AT);
constexpr uint32_t kVecPerWg = 32, kAdamVec = (770 + kVecPerWg - 1) / kVecPerWg;

__global__ void __launch_bounds__(256) k_adam(AdamArgs a) {
  __shared__ uint16_t tile[AT][AT + 2];
  const uint32_t t = threadIdx.x, b = blockIdx.x;
  if (b < kAdamW1 + kAdamW2) {
    const bool w1 = b < kAdamW1;
    const uint32_t tb = w1 ? b : b - kAdamW1;
    const uint32_t i0 = (tb / (256 / AT)) * AT, o0 = (tb % (256 / AT)) * AT;
    const uint32_t il = t / AT, ol = t % AT;
    const uint32_t q = (i0 + il) * 256 + o0 + ol;  // index within the matrix ([in][out])
    const uint32_t p = (w1 ? MI_MLP_OFF_W1 : MI_MLP_OFF_W2) + q;
    const float w = a.w[p], m = a.m[p], v = a.v[p];
    const float g = sum_partials(a.P + (w1 ? 0u : 32768u) + q, kWParams, a.splits);
    const uint16_t h = to_bf(adam_step(a, p, g, w, m, v));
    if (!w1) a.W2r[q] = h;
    tile[ol][il] = h;
    lds_barrier();
    const uint32_t ro = t / AT, ri = t % AT;  // row (out) and column (in) of the transposed copy
    (w1 ? a.W1t : a.W2t)[(o0 + ro) * (w1 ? 128u : 256u) + i0 + ri] = tile[ro][ri];
    return;
  }
  __shared__ float eighth[8][kVecPerWg];
  const uint32_t sl = t % kVecPerWg, e = t / kVecPerWg;  // slot in the workgroup, eighth
  const uint32_t slot = (b - kAdamW1 - kAdamW2) * kVecPerWg + sl;  // b1 0-255, b2 256-511,
  const uint32_t per = (a.blocks + 7u) / 8u;                      // w3 512-767, b3 768, loss 769
  const uint32_t lo = min(e * per, a.blocks), hi = min(lo + per, a.blocks);
  eighth[e][sl] = slot <= 769u ? sum_partials(a.pV + (size_t)lo * kVecSlots + slot, kVecSlots, hi - lo)
                               : 0.f;
  lds_barrier();
  if (e != 0u || slot > 769u) return;
  float g = 0.f;
#pragma unroll
  for (int k = 0; k < 8; k++) g += eighth[k][sl];
  if (slot == 769u) {
    if (a.loss_out) *a.loss_out = g;
    return;
  }
  const uint32_t p = slot < 256u ? MI_MLP_OFF_B1 + slot : MI_MLP_OFF_B2 + (slot - 256u);
  adam_step(a, p, g, a.w[p], a.m[p], a.v[p]);
}

__global__ void k_zero(float *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

}  // namespace

hipError_t forward(const mi_pos *pos, uint32_t n, const uint16_t *W1t, const float *b1,
                   const uint16_t *W2t, const float *b2, const float *w3, const float *b3, float *y,
                   hipStream_t stream) {
  hipLaunchKernelGGL(k_forward, dim3((n + FR - 1) / FR), dim3(256), 0, stream, pos, n, W1t, b1,
                     W2t, b2, w3, b3, y);
  return hipGetLastError();
}

hipError_t train_rows(const TrainArgs &a, uint32_t rows, hipStream_t stream) {
  if (rows % KR || rows == 0 || rows > a.ld || a.n == 0 || a.n > rows) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_train_rows, dim3(rows / KR), dim3(512), 0, stream, a);
  return hipGetLastError();
}

hipError_t wgrad_adam(const AdamArgs &a, hipStream_t stream) {
  if (a.rows % KR || a.rows > a.ld || a.blocks * KR != a.rows ||
      a.splits != (a.rows + kWSplit - 1) / kWSplit)
    return hipErrorInvalidValue;
  if (a.splits) hipLaunchKernelGGL(k_wgrad_part, dim3(kWTiles, a.splits), dim3(256), 0, stream, a);
  hipLaunchKernelGGL(k_adam, dim3(kAdamW1 + kAdamW2 + kAdamVec), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t zero(float *p, uint32_t n, hipStream_t stream) {
  hipLaunchKernelGGL(k_zero, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace mlp
}  // namespace gpuai
