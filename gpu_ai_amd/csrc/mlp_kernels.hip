// mlp_kernels.hip -- gfx950 kernels of the 128-256-256-1 value network.
//
// k_gemm: C[M][N] = A[M][K] B[K][N] on v_mfma_f32_16x16x32_bf16.  A 256-thread workgroup owns a
// 64x64 tile of C; its four waves own 32x32 quadrants (2x2 MFMA tiles of 16x16).  Each 32-deep
// slice of K is staged through LDS as BF16 with K contiguous (A as [m][k], B as [n][k]), so every
// lane feeds the MFMA with one 16-byte LDS read per operand: lane l supplies row l%16 of A and
// column l%16 of B for eight K values of its lane group l/16 (the same eight for A and B, which
// is all the dot product needs), and receives D[4*(l/16)+v][l%16], v = 0..3.  A and B are read
// from global memory through arbitrary (row, column) strides, which gives the transposed
// operands of the backward pass without transposing in memory.  Epilogues fuse bias, ReLU, the
// ReLU mask of the backward pass, the BF16 store of activations and split-K partial sums.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_kernels.h"

namespace gpuai {
namespace mlp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32, LDSK = 40;  // LDS rows padded to 40 elements

__device__ __forceinline__ uint16_t to_bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// One thread's share of a 64 x 32 slice of a strided BF16 operand X(r, k) = X[r*sr + k*sk]: eight
// values, with K contiguous (sk == 1: row t/4, k 8(t%4)..+7) or rows contiguous (sr == 1: k t/8,
// rows 8(t%8)..+7), both one 16-byte load; anything else element by element.
__device__ __forceinline__ uint4 load_slice(const uint16_t *__restrict__ X, long sr, long sk,
                                            int r0, int k0) {
  const int t = threadIdx.x;
  if (sk == 1) {
    return *(const uint4 *)(X + (long)(r0 + (t >> 2)) * sr + (k0 + (t & 3) * 8));
  } else if (sr == 1) {
    return *(const uint4 *)(X + (long)(r0 + (t & 7) * 8) + (long)(k0 + (t >> 3)) * sk);
  }
  uint16_t e[8];
  const int kk = t >> 3, rr = (t & 7) * 8;
#pragma unroll
  for (int j = 0; j < 8; j++) e[j] = X[(long)(r0 + rr + j) * sr + (long)(k0 + kk) * sk];
  uint4 v;
  v.x = e[0] | ((uint32_t)e[1] << 16);
  v.y = e[2] | ((uint32_t)e[3] << 16);
  v.z = e[4] | ((uint32_t)e[5] << 16);
  v.w = e[6] | ((uint32_t)e[7] << 16);
  return v;
}

// Stores a thread's share into the LDS tile S[row][k] (K contiguous, rows padded to LDSK).
__device__ __forceinline__ void store_slice(uint16_t (*S)[LDSK], long sk, const uint4 &v) {
  const int t = threadIdx.x;
  if (sk == 1) {
    *(uint4 *)&S[t >> 2][(t & 3) * 8] = v;
    return;
  }
  const int kk = t >> 3, rr = (t & 7) * 8;
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int j = 0; j < 4; j++) {
    S[rr + 2 * j][kk] = (uint16_t)(w[j] & 0xFFFFu);
    S[rr + 2 * j + 1][kk] = (uint16_t)(w[j] >> 16);
  }
}

// C = A B over one 64x64 tile and one K split; the next 32-deep slice is loaded into registers
// while the current one is multiplied.
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[BM][LDSK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN][LDSK];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int kbeg = blockIdx.z * g.k_per_split;
  const int kend = kbeg + g.k_per_split;
  const uint16_t *A = (const uint16_t *)g.A;
  const uint16_t *B = (const uint16_t *)g.B;
  f32x4 acc[2][2];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int c = 0; c < 2; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  uint4 ra = load_slice(A, g.sAm, g.sAk, m0, kbeg);
  uint4 rb = load_slice(B, g.sBn, g.sBk, n0, kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    store_slice(As, g.sAk, ra);
    store_slice(Bs, g.sBk, rb);
    __syncthreads();
    if (k0 + BK < kend) {  // in flight during the MFMAs below
      ra = load_slice(A, g.sAm, g.sAk, m0, k0 + BK);
      rb = load_slice(B, g.sBn, g.sBk, n0, k0 + BK);
    }
    bf16x8 a[2], b[2];
#pragma unroll
    for (int r = 0; r < 2; r++)
      a[r] = *(const bf16x8 *)&As[wr * 32 + r * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
    for (int c = 0; c < 2; c++)
      b[c] = *(const bf16x8 *)&Bs[wc * 32 + c * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
      for (int c = 0; c < 2; c++)
        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b[c], acc[r][c], 0, 0, 0);
    __syncthreads();
  }
  float *C = g.C ? g.C + (long)blockIdx.z * g.split_stride : nullptr;
#pragma unroll
  for (int r = 0; r < 2; r++) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int n = n0 + wc * 32 + c * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = m0 + wr * 32 + r * 16 + 4 * (lane >> 4) + v;
        float x = acc[r][c][v];
        if (g.bias) x += g.bias[n];
        if (g.relu) x = x > 0.f ? x : 0.f;
        if (g.mask) x = from_bf(g.mask[(long)m * g.ldm + n]) > 0.f ? x : 0.f;
        if (g.row_valid && m >= g.row_valid) x = 0.f;
        if (C) C[(long)m * g.ldc + n] = x;
        if (g.Cbf) g.Cbf[(long)m * g.ldc + n] = to_bf(x);
      }
    }
  }
}

// Positions -> BF16 input planes in the mover's frame; rows at or past n (padding) are zero.
__global__ void __launch_bounds__(256) k_encode(const mi_pos *__restrict__ pos,
                                               const uint32_t *__restrict__ idx, uint32_t n,
                                               uint32_t rows, uint16_t *__restrict__ X) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;  // one thread per (row, plane)
  const uint32_t row = t >> 2, plane = t & 3;
  if (row >= rows) return;
  uint32_t bits = 0;
  if (row < n) {
    const mi_pos p = pos[idx ? idx[row] : row];
    uint32_t own = p.turn ? __builtin_bitreverse32(p.p2) : p.p1;
    uint32_t opp = p.turn ? __builtin_bitreverse32(p.p1) : p.p2;
    uint32_t kings = p.turn ? __builtin_bitreverse32(p.kings) : p.kings;
    kings &= own | opp;
    bits = plane == 0 ? (own & ~kings) : plane == 1 ? (own & kings)
         : plane == 2 ? (opp & ~kings) : (opp & kings);
  }
  uint16_t *dst = X + (size_t)row * 128 + plane * 32;
#pragma unroll
  for (int s = 0; s < 32; s++) dst[s] = ((bits >> s) & 1u) ? (uint16_t)0x3F80u : (uint16_t)0u;
}

// Inference in one launch: encoding, both hidden layers and the output layer for FR rows per
// workgroup, with the activations in LDS.  The arithmetic is that of k_encode and k_gemm (bias,
// ReLU, BF16 rounding of H1 and H2; the same 32-deep K slices accumulated in the same order by
// the same MFMA), and the output layer sums four columns per lane and then over the wave with a
// butterfly, so the values equal those of the layer-by-layer forward bit for bit (checked at
// batches of 1 to 4096 rows).  Wave w computes output columns 64w..64w+63 of
// each layer as four 16x16 tiles; the B operands (W1t, W2t: [out][in], in contiguous) are read
// from global memory (they stay in L2), one 16-byte load per lane per tile and slice.
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
    if (r0 + row < n) {
      const mi_pos p = pos[r0 + row];
      uint32_t own = p.turn ? __builtin_bitreverse32(p.p2) : p.p1;
      uint32_t opp = p.turn ? __builtin_bitreverse32(p.p1) : p.p2;
      uint32_t kings = p.turn ? __builtin_bitreverse32(p.kings) : p.kings;
      kings &= own | opp;
      bits = plane == 0 ? (own & ~kings) : plane == 1 ? (own & kings)
           : plane == 2 ? (opp & ~kings) : (opp & kings);
    }
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

// Output layer with its loss and gradients, over 64-row blocks (four waves of 16 rows; lane l
// holds columns 4l..4l+3).  Per row: y = tanh(H2 . w3 + b3), e = y - t, dz = 2 e (1 - y^2) / n
// (0 on padding rows) and dH2 = dz w3 masked by H2 > 0 (BF16).  Per block, into the partial
// slots of block b (part + b * kVecSlots): b2 (256 + c) = sum of dH2, w3 (512 + c) = sum of
// dz H2, b3 (768) = sum of dz, and the loss (769) = sum of e^2 / n.  No atomics: the partials
// are summed in a fixed order by k_adam.
__global__ void __launch_bounds__(256) k_head_train(const uint16_t *__restrict__ H2,
                                                   const float *__restrict__ w3, const float *b3,
                                                   const float *__restrict__ targets,
                                                   const uint32_t *__restrict__ tidx, uint32_t n,
                                                   float *__restrict__ dz,
                                                   uint16_t *__restrict__ dH2,
                                                   float *__restrict__ part) {
  __shared__ float red_b2[4][256], red_w3[4][256], red_s[4][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float wl[4], acc_b2[4] = {0.f, 0.f, 0.f, 0.f}, acc_w3[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < 4; c++) wl[c] = w3[lane * 4 + c];
  const float bias = b3[0];
  float acc_b3 = 0.f, acc_loss = 0.f;
  const uint32_t row0 = blockIdx.x * 64u + wv * 16u;
  // the wave's 16 targets, fetched together up front (lane i holds row i's) instead of one
  // dependent pair of loads per row
  float tgt = 0.f;
  if (lane < 16 && row0 + lane < n) tgt = targets[tidx ? tidx[row0 + lane] : row0 + lane];
  // the 16 rows' activations, loaded together (4 BF16 per lane and row, one 8-byte load each);
  // the row loop is unrolled so that they stay in registers
  uint2 hrow[16];
#pragma unroll
  for (uint32_t i = 0; i < 16u; i++)
    hrow[i] = *(const uint2 *)(H2 + (size_t)(row0 + i) * 256 + lane * 4);
#pragma unroll
  for (uint32_t i = 0; i < 16u; i++) {
    const uint32_t row = row0 + i;
    const float target = __shfl(tgt, (int)i);
    const uint2 h = hrow[i];
    const float hv[4] = {from_bf((uint16_t)(h.x & 0xFFFFu)), from_bf((uint16_t)(h.x >> 16)),
                         from_bf((uint16_t)(h.y & 0xFFFFu)), from_bf((uint16_t)(h.y >> 16))};
    float dot = 0.f;
#pragma unroll
    for (int c = 0; c < 4; c++) dot += hv[c] * wl[c];
    for (int off = 32; off >= 1; off >>= 1) dot += __shfl_xor(dot, off);
    const float yy = tanhf(dot + bias);
    float d = 0.f;
    if (row < n) {
      const float e = yy - target;
      d = 2.f * e * (1.f - yy * yy) / (float)n;
      acc_loss += e * e / (float)n;
    }
    if (lane == 0) dz[row] = d;
    acc_b3 += d;
    uint16_t gb[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      gb[c] = to_bf(hv[c] > 0.f ? d * wl[c] : 0.f);
      acc_b2[c] += from_bf(gb[c]);
      acc_w3[c] += d * hv[c];
    }
    *(uint2 *)(dH2 + (size_t)row * 256 + lane * 4) =
        uint2{gb[0] | ((uint32_t)gb[1] << 16), gb[2] | ((uint32_t)gb[3] << 16)};
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    red_b2[wv][lane * 4 + c] = acc_b2[c];
    red_w3[wv][lane * 4 + c] = acc_w3[c];
  }
  if (lane == 0) {
    red_s[wv][0] = acc_b3;
    red_s[wv][1] = acc_loss;
  }
  __syncthreads();
  float *pb = part + (size_t)blockIdx.x * kVecSlots;
  const int t = threadIdx.x;
  pb[256 + t] = red_b2[0][t] + red_b2[1][t] + red_b2[2][t] + red_b2[3][t];
  pb[512 + t] = red_w3[0][t] + red_w3[1][t] + red_w3[2][t] + red_w3[3][t];
  if (t == 0) {
    pb[768] = red_s[0][0] + red_s[1][0] + red_s[2][0] + red_s[3][0];
    pb[769] = red_s[0][1] + red_s[1][1] + red_s[2][1] + red_s[3][1];
  }
}

// Partial column sums: block (x, y) writes part[y * part_stride + col] = sum over rows
// [64 y, 64 y + 64) of weight[row] * src[row][col] (weight 1 when null), for 64 columns of a BF16
// or FP32 source with `cols` columns.  The partials are summed in a fixed order by k_adam, so
// the reduction is deterministic.
template <typename T>
__global__ void __launch_bounds__(256) k_colsum(const T *__restrict__ src, uint32_t rows,
                                               uint32_t cols, const float *__restrict__ weight,
                                               float *__restrict__ part, uint32_t part_stride) {
  __shared__ float red[4][64];
  const int c = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const uint32_t col = blockIdx.x * 64 + c;
  const uint32_t r0 = blockIdx.y * 64, r1 = min(r0 + 64u, rows);
  float s = 0.f;
  if (col < cols) {
#pragma unroll 4
    for (uint32_t r = r0 + grp; r < r1; r += 4) {
      float v;
      if constexpr (sizeof(T) == 2) v = from_bf((uint16_t)src[(size_t)r * cols + col]);
      else v = (float)src[(size_t)r * cols + col];
      s += weight ? weight[r] * v : v;
    }
  }
  red[grp][c] = s;
  __syncthreads();
  if (grp == 0 && col < cols)
    part[(size_t)blockIdx.y * part_stride + col] = red[0][c] + red[1][c] + red[2][c] + red[3][c];
}

// Adam over every parameter.  Gradients are sums of partials in a fixed order: W1 and W2 over
// their split-K partials, b1, b2, w3 and b3 over the column-sum partials (slots 0-255 b1, then
// b2, w3, b3 at 256-768, in parameter order).  The summed gradient is written to `grads` (for
// inspection).  Refreshes the BF16 copies: W1t[o][i], W2t[o][i] (forward B operands) and
// W2r[i][o] (backward).
// Adam over 16x16 tiles of W1 ([128][256]) and W2 ([256][256]), then the 769 vector parameters
// (b1, b2, w3, b3; slot v of the column-sum partials).  Each parameter's gradient is the sum of
// its partials in split order; the BF16 copies are refreshed, the transposed ones (W1t, W2t:
// [out][in]) through an LDS tile so that their stores are contiguous.
constexpr uint32_t AT = 16, kAdamW1 = (128 / AT) * (256 / AT), kAdamW2 = (256 / AT) * (256 / AT);
constexpr uint32_t kAdamVec = (769 + 255) / 256;

__device__ __forceinline__ float adam_step(const AdamArgs &a, uint32_t p, float g) {
  a.grads[p] = g;
  float w = a.w[p];
  if (a.update) {
    const float m = a.beta1 * a.m[p] + a.omb1 * g;
    const float v = a.beta2 * a.v[p] + a.omb2 * g * g;
    a.m[p] = m;
    a.v[p] = v;
    w -= a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
    a.w[p] = w;
  }
  return w;
}

__global__ void __launch_bounds__(256) k_adam(AdamArgs a) {
  __shared__ uint16_t tile[AT][AT + 2];
  const uint32_t t = threadIdx.x, b = blockIdx.x;
  if (b < kAdamW1 + kAdamW2) {
    const bool w1 = b < kAdamW1;
    const uint32_t tb = w1 ? b : b - kAdamW1;
    const uint32_t i0 = (tb / (256 / AT)) * AT, o0 = (tb % (256 / AT)) * AT;
    const uint32_t il = t / AT, ol = t % AT;
    const uint32_t q = (i0 + il) * 256 + o0 + ol;  // index within the matrix ([in][out])
    const float *pW = w1 ? a.pW1 : a.pW2;
    const size_t stride = w1 ? 32768 : 65536;
    float g = 0.f;
#pragma unroll 8
    for (uint32_t s = 0; s < a.splits; s++) g += pW[(size_t)s * stride + q];
    const uint16_t h = to_bf(adam_step(a, (w1 ? MI_MLP_OFF_W1 : MI_MLP_OFF_W2) + q, g));
    if (!w1) a.W2r[q] = h;
    tile[ol][il] = h;
    __syncthreads();
    const uint32_t ro = t / AT, ri = t % AT;  // row (out) and column (in) of the transposed copy
    (w1 ? a.W1t : a.W2t)[(o0 + ro) * (w1 ? 128u : 256u) + i0 + ri] = tile[ro][ri];
    return;
  }
  if (b == kAdamW1 + kAdamW2 && t < 64u && a.loss_out) {
    // the batch loss: the sum of the head's per-block partials in block order; the first wave
    // loads 64 partials at a time and every lane adds them in order from the lanes that hold
    // them
    float l = 0.f;
    for (uint32_t base = 0; base < a.vsplits; base += 64u) {
      const float part = base + t < a.vsplits ? a.pV[(size_t)(base + t) * kVecSlots + 769] : 0.f;
      const uint32_t cnt = min(64u, a.vsplits - base);
      for (uint32_t j = 0; j < cnt; j++) l += __shfl(part, (int)j);
    }
    if (t == 0u) *a.loss_out = l;
  }
  const uint32_t v = (b - kAdamW1 - kAdamW2) * 256 + t;  // vector slot: b1 0-255, b2 256-511,
  if (v >= 769u) return;                                 // w3 512-767, b3 768
  float g = 0.f;
#pragma unroll 8
  for (uint32_t s = 0; s < a.vsplits; s++) g += a.pV[(size_t)s * kVecSlots + v];
  adam_step(a, v < 256u ? MI_MLP_OFF_B1 + v : MI_MLP_OFF_B2 + (v - 256u), g);
}

__global__ void k_zero(float *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

}  // namespace

hipError_t gemm(const GemmArgs &g, int splits, bool a_bf16, bool b_bf16, hipStream_t stream) {
  if (g.M % BM || g.N % BN || g.k_per_split % BK) return hipErrorInvalidValue;
  if (!a_bf16 || !b_bf16) return hipErrorInvalidValue;  // BF16 operands only
  dim3 grid(g.N / BN, g.M / BM, splits), block(256);
  hipLaunchKernelGGL(k_gemm, grid, block, 0, stream, g);
  return hipGetLastError();
}

hipError_t encode(const mi_pos *pos, const uint32_t *idx, uint32_t n, uint32_t rows, uint16_t *X,
                  hipStream_t stream) {
  const uint32_t threads = rows * 4;
  hipLaunchKernelGGL(k_encode, dim3((threads + 255) / 256), dim3(256), 0, stream, pos, idx, n, rows, X);
  return hipGetLastError();
}

hipError_t forward(const mi_pos *pos, uint32_t n, const uint16_t *W1t, const float *b1,
                   const uint16_t *W2t, const float *b2, const float *w3, const float *b3, float *y,
                   hipStream_t stream) {
  hipLaunchKernelGGL(k_forward, dim3((n + FR - 1) / FR), dim3(256), 0, stream, pos, n, W1t, b1,
                     W2t, b2, w3, b3, y);
  return hipGetLastError();
}

hipError_t head_train(const uint16_t *H2, const float *w3, const float *b3, const float *targets,
                      const uint32_t *tidx, uint32_t n, uint32_t rows, float *dz, uint16_t *dH2,
                      float *part, hipStream_t stream) {
  if (rows % 64u) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_head_train, dim3(rows / 64), dim3(256), 0, stream, H2, w3, b3, targets,
                     tidx, n, dz, dH2, part);
  return hipGetLastError();
}

hipError_t colsum_bf16(const uint16_t *src, uint32_t rows, uint32_t cols, const float *weight,
                       float *part, hipStream_t stream) {
  hipLaunchKernelGGL(k_colsum<uint16_t>, dim3((cols + 63) / 64, (rows + 63) / 64), dim3(256), 0,
                     stream, src, rows, cols, weight, part, (uint32_t)kVecSlots);
  return hipGetLastError();
}

hipError_t colsum_f32(const float *src, uint32_t rows, uint32_t cols, const float *weight,
                      float *part, hipStream_t stream) {
  hipLaunchKernelGGL(k_colsum<float>, dim3((cols + 63) / 64, (rows + 63) / 64), dim3(256), 0,
                     stream, src, rows, cols, weight, part, (uint32_t)kVecSlots);
  return hipGetLastError();
}

hipError_t adam(const AdamArgs &a, hipStream_t stream) {
  hipLaunchKernelGGL(k_adam, dim3(kAdamW1 + kAdamW2 + kAdamVec), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t zero(float *p, uint32_t n, hipStream_t stream) {
  hipLaunchKernelGGL(k_zero, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace mlp
}  // namespace gpuai
