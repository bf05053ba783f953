// mlp.cpp -- the value network's C ABI (include/gpuai/mi_mlp.h): buffers, the sequence of
// kernels of a forward pass, of a backward pass and of an Adam step, device-resident epochs and
// checkpoints.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "checkers.h"
#include "ctx.h"
#include "gpuai/mi_mlp.h"
#include "mlp_kernels.h"

using gpuai::DevBuf;
using gpuai::ctxutil::ensure;
using gpuai::ctxutil::fail;
using gpuai::ctxutil::hip_fail;
namespace K = gpuai::mlp;

struct mi_mlp {
  mi_ctx *ctx = nullptr;
  uint32_t max_rows = 0;
  uint32_t step = 0;
  double beta1 = 0.9, beta2 = 0.999, eps = 1e-8;  // rounded to float where a kernel takes them
  // parameters, Adam moments, gradients (FP32) and BF16 operand copies
  float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  uint16_t *W1t = nullptr, *W2t = nullptr, *W2r = nullptr;
  // one batch: network outputs, targets, loss; the transposed activations of a training step
  // ([feature][max_rows]) and its per 32-row block partials of the vector gradients
  float *y = nullptr, *targets = nullptr, *loss = nullptr;
  uint16_t *XT = nullptr, *H1T = nullptr, *dH2T = nullptr, *dH1T = nullptr;
  float *P = nullptr;  // per 256-row split partials of dW1 and dW2
  float *pV = nullptr;
  mi_pos *pos = nullptr;
  // pinned host staging of one batch: positions (max_rows), then values and targets (max_rows
  // floats each); copies from and to the caller's pageable buffers go through it, so the DMA
  // transfers need no driver-side staging
  void *h_stage = nullptr;
  // device-resident training set
  DevBuf ds_pos, ds_t, ds_perm, ds_loss;
  uint32_t ds_n = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<void *> allocs;
};

namespace {

uint32_t round64(uint32_t n) { return (n + 63u) & ~63u; }

// Makes the network's device current on the calling thread (as every playout entry point does).
int use_device(mi_mlp *n) {
  if (hipSetDevice(n->ctx->device) != hipSuccess)
    return fail(n->ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  return MI_OK;
}

mi_pos *stage_pos(mi_mlp *n) { return (mi_pos *)n->h_stage; }
float *stage_values(mi_mlp *n) { return (float *)((char *)n->h_stage + sizeof(mi_pos) * n->max_rows); }
float *stage_targets(mi_mlp *n) { return stage_values(n) + n->max_rows; }

K::AdamArgs adam_args(mi_mlp *n, uint32_t rows, float lr, float c1, float c2, int update,
                      float *loss_out = nullptr) {
  K::AdamArgs a = {};
  a.loss_out = loss_out;
  a.w = n->w;
  a.m = n->m;
  a.v = n->v;
  a.grads = n->g;
  a.XT = n->XT;
  a.H1T = n->H1T;
  a.dH2T = n->dH2T;
  a.dH1T = n->dH1T;
  a.rows = rows;
  a.ld = n->max_rows;
  a.P = n->P;
  a.splits = (rows + K::kWSplitRows - 1) / K::kWSplitRows;
  a.pV = n->pV;
  a.blocks = rows / K::kTrainRows;
  a.W1t = n->W1t;
  a.W2t = n->W2t;
  a.W2r = n->W2r;
  a.lr = lr;
  a.beta1 = (float)n->beta1;
  a.beta2 = (float)n->beta2;
  a.eps = (float)n->eps;
  a.c1 = c1;
  a.c2 = c2;
  a.omb1 = (float)(1.0 - n->beta1);  // formed in double precision, as Adam's usual
  a.omb2 = (float)(1.0 - n->beta2);  // implementations do
  a.update = update;
  return a;
}

template <class T>
hipError_t dev_alloc(mi_mlp *n, T **p, size_t count) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e == hipSuccess) {
    n->allocs.push_back(q);
    *p = (T *)q;
  }
  return e;
}

// He-uniform initialisation: U(-a, a), a = sqrt(6 / fan_in), from Philox key (seed, tensor),
// counter (i >> 2, 0, 3, 0), word i & 3, mapped to [0, 1) by its top 24 bits.
void init_params(uint32_t seed, std::vector<float> &w) {
  w.assign(MI_MLP_PARAMS, 0.f);
  struct T { uint32_t off, count, fan_in; } tensors[] = {
      {MI_MLP_OFF_W1, 128 * 256, 128}, {MI_MLP_OFF_W2, 256 * 256, 256}, {MI_MLP_OFF_W3, 256, 256}};
  for (uint32_t t = 0; t < 3; t++) {
    const float a = std::sqrt(6.f / (float)tensors[t].fan_in);
    uint32_t words[4];
    for (uint32_t i = 0; i < tensors[t].count; i++) {
      if ((i & 3) == 0) ck::philox(i >> 2, 0u, ck::TAG_MLP, 0u, seed, t, words);
      const float u = (float)(words[i & 3] >> 8) * (1.f / 16777216.f);
      w[tensors[t].off + i] = (2.f * u - 1.f) * a;
    }
  }
}

int upload_params(mi_mlp *n, const float *w) {
  mi_ctx *ctx = n->ctx;
  MI_HIP(ctx, hipMemcpyAsync(n->w, w, sizeof(float) * MI_MLP_PARAMS, hipMemcpyHostToDevice, ctx->stream));
  MI_HIP(ctx, K::zero(n->m, MI_MLP_PARAMS, ctx->stream));
  MI_HIP(ctx, K::zero(n->v, MI_MLP_PARAMS, ctx->stream));
  MI_HIP(ctx, K::zero(n->g, MI_MLP_PARAMS, ctx->stream));
  // refresh the BF16 copies without updating (no partials: every gradient reads as 0)
  MI_HIP(ctx, K::wgrad_adam(adam_args(n, 0, 0.f, 1.f, 1.f, 0), ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  n->step = 0;
  return MI_OK;
}

// One training step over `rows` rows (count valid): k_train_rows' forward and backward pass to
// the transposed activations and vector partials, then the weight-gradient partials and Adam
// (k_wgrad_part, k_adam; when update == 0, only the gradient sums); loss_out (device, may be
// null) receives the batch loss.  Row r is
// pos[idx ? idx[r] : r] with target targets[idx ? idx[r] : r].
int train_rows(mi_mlp *n, const mi_pos *pos, const uint32_t *idx, const float *targets,
               uint32_t count, uint32_t rows, float lr, int update, float *loss_out) {
  mi_ctx *ctx = n->ctx;
  K::TrainArgs ta = {};
  ta.pos = pos;
  ta.idx = idx;
  ta.targets = targets;
  ta.n = count;
  ta.ld = n->max_rows;
  ta.W1t = n->W1t;
  ta.W2t = n->W2t;
  ta.W2r = n->W2r;
  ta.b1 = n->w + MI_MLP_OFF_B1;
  ta.b2 = n->w + MI_MLP_OFF_B2;
  ta.w3 = n->w + MI_MLP_OFF_W3;
  ta.b3 = n->w + MI_MLP_OFF_B3;
  ta.XT = n->XT;
  ta.H1T = n->H1T;
  ta.dH2T = n->dH2T;
  ta.dH1T = n->dH1T;
  ta.pV = n->pV;
  MI_HIP(ctx, K::train_rows(ta, rows, ctx->stream));
  // the bias corrections are formed in double precision, as Adam's usual implementations do
  float c1 = 1.f, c2 = 1.f;
  if (update) {
    n->step++;
    c1 = (float)(1.0 / (1.0 - std::pow(n->beta1, (double)n->step)));
    c2 = (float)(1.0 / (1.0 - std::pow(n->beta2, (double)n->step)));
  }
  MI_HIP(ctx, K::wgrad_adam(adam_args(n, rows, lr, c1, c2, update, loss_out), ctx->stream));
  return MI_OK;
}

int check_batch(mi_mlp *n, uint32_t count, const void *pos) {
  if (count == 0) return MI_OK;
  if (!pos) return fail(n->ctx, MI_ERR_INVALID_ARGUMENT, "null positions");
  if (round64(count) > n->max_rows)
    return fail(n->ctx, MI_ERR_INVALID_ARGUMENT, "batch larger than the network's max_batch");
  return MI_OK;
}

// Uploads the positions of a batch through the pinned staging buffer.
int upload_positions(mi_mlp *n, const mi_pos *pos, uint32_t count) {
  mi_ctx *ctx = n->ctx;
  mi_pos *hp = stage_pos(n);
  std::memcpy(hp, pos, sizeof(mi_pos) * count);
  MI_HIP(ctx, hipMemcpyAsync(n->pos, hp, sizeof(mi_pos) * count, hipMemcpyHostToDevice, ctx->stream));
  return MI_OK;
}

// Uploads a batch (positions, optional targets); returns the padded row count.
int stage_batch(mi_mlp *n, const mi_pos *pos, const float *targets, uint32_t count,
                uint32_t &rows) {
  mi_ctx *ctx = n->ctx;
  rows = round64(count);
  int rc;
  if ((rc = upload_positions(n, pos, count))) return rc;
  if (targets) {
    float *ht = stage_targets(n);
    std::memcpy(ht, targets, sizeof(float) * count);
    MI_HIP(ctx, hipMemcpyAsync(n->targets, ht, sizeof(float) * count, hipMemcpyHostToDevice,
                               ctx->stream));
  }
  return MI_OK;
}

}  // namespace

extern "C" {

int mi_mlp_create(mi_ctx *ctx, uint32_t max_batch, uint32_t seed, mi_mlp **out) {
  if (!ctx || !out || max_batch == 0) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  *out = nullptr;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  mi_mlp *n = new mi_mlp();
  n->ctx = ctx;
  n->max_rows = round64(max_batch);
  const size_t R = n->max_rows, P = MI_MLP_PARAMS;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = dev_alloc(n, &n->w, P);
  if (e == hipSuccess) e = dev_alloc(n, &n->m, P);
  if (e == hipSuccess) e = dev_alloc(n, &n->v, P);
  if (e == hipSuccess) e = dev_alloc(n, &n->g, P);
  if (e == hipSuccess) e = dev_alloc(n, &n->W1t, (size_t)128 * 256);
  if (e == hipSuccess) e = dev_alloc(n, &n->W2t, (size_t)256 * 256);
  if (e == hipSuccess) e = dev_alloc(n, &n->W2r, (size_t)256 * 256);
  if (e == hipSuccess) e = dev_alloc(n, &n->y, R);
  if (e == hipSuccess) e = dev_alloc(n, &n->targets, R);
  if (e == hipSuccess) e = dev_alloc(n, &n->loss, 1);
  if (e == hipSuccess) e = dev_alloc(n, &n->XT, 128 * R);
  if (e == hipSuccess) e = dev_alloc(n, &n->H1T, 256 * R);
  if (e == hipSuccess) e = dev_alloc(n, &n->dH2T, 256 * R);
  if (e == hipSuccess) e = dev_alloc(n, &n->dH1T, 256 * R);
  if (e == hipSuccess)
    e = dev_alloc(n, &n->P, ((R + K::kWSplitRows - 1) / K::kWSplitRows) * (size_t)(32768 + 65536));
  if (e == hipSuccess) e = dev_alloc(n, &n->pV, (R / K::kTrainRows) * K::kVecSlots);
  if (e == hipSuccess) e = dev_alloc(n, &n->pos, R);
  if (e == hipSuccess)
    e = hipHostMalloc(&n->h_stage, R * (sizeof(mi_pos) + 2 * sizeof(float)), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreate(&n->e0);
  if (e == hipSuccess) e = hipEventCreate(&n->e1);
  int rc = MI_OK;
  if (e != hipSuccess) rc = hip_fail(ctx, e, "mi_mlp_create");
  std::vector<float> w;
  init_params(seed, w);
  if (!rc) rc = upload_params(n, w.data());
  if (rc) {
    for (void *p : n->allocs) (void)hipFree(p);
    if (n->h_stage) (void)hipHostFree(n->h_stage);
    if (n->e0) (void)hipEventDestroy(n->e0);
    if (n->e1) (void)hipEventDestroy(n->e1);
    delete n;
    return rc;
  }
  *out = n;
  return MI_OK;
}

int mi_mlp_destroy(mi_mlp *n) {
  if (!n) return MI_OK;
  std::lock_guard<std::mutex> lock(n->ctx->mu);
  (void)hipSetDevice(n->ctx->device);
  (void)hipStreamSynchronize(n->ctx->stream);
  for (void *p : n->allocs) (void)hipFree(p);
  if (n->h_stage) (void)hipHostFree(n->h_stage);
  gpuai::DevBuf *bufs[] = {&n->ds_pos, &n->ds_t, &n->ds_perm, &n->ds_loss};
  for (auto *b : bufs)
    if (b->ptr) (void)hipFree(b->ptr);
  if (n->e0) (void)hipEventDestroy(n->e0);
  if (n->e1) (void)hipEventDestroy(n->e1);
  delete n;
  return MI_OK;
}

int mi_mlp_get_params(mi_mlp *n, float *params) {
  if (!n || !params) return MI_ERR_INVALID_ARGUMENT;
  mi_ctx *ctx = n->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (int rc = use_device(n)) return rc;
  MI_HIP(ctx, hipMemcpyAsync(params, n->w, sizeof(float) * MI_MLP_PARAMS, hipMemcpyDeviceToHost,
                             ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MI_OK;
}

int mi_mlp_set_params(mi_mlp *n, const float *params) {
  if (!n || !params) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(n->ctx->mu);
  n->ctx->last_error.clear();
  if (int rc = use_device(n)) return rc;
  return upload_params(n, params);
}

int mi_mlp_forward(mi_mlp *n, const mi_pos *pos, uint32_t count, float *values) {
  if (!n) return MI_ERR_INVALID_ARGUMENT;
  mi_ctx *ctx = n->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  int rc;
  if ((rc = check_batch(n, count, pos))) return rc;
  if (count == 0) return MI_OK;
  if (!values) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "null values");
  if ((rc = use_device(n))) return rc;
  // one launch: encoding, both hidden layers and the output layer (k_forward)
  if ((rc = upload_positions(n, pos, count))) return rc;
  MI_HIP(ctx, K::forward(n->pos, count, n->W1t, n->w + MI_MLP_OFF_B1, n->W2t, n->w + MI_MLP_OFF_B2,
                         n->w + MI_MLP_OFF_W3, n->w + MI_MLP_OFF_B3, n->y, ctx->stream));
  float *hv = stage_values(n);
  MI_HIP(ctx, hipMemcpyAsync(hv, n->y, sizeof(float) * count, hipMemcpyDeviceToHost, ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(values, hv, sizeof(float) * count);
  return MI_OK;
}

static int grad_common(mi_mlp *n, const mi_pos *pos, const float *targets, uint32_t count,
                       float lr, int update, float *grads, float *loss) {
  mi_ctx *ctx = n->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  int rc;
  if ((rc = check_batch(n, count, pos))) return rc;
  if (count == 0 || !targets) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "empty batch or null targets");
  if ((rc = use_device(n))) return rc;
  uint32_t rows;
  if ((rc = stage_batch(n, pos, targets, count, rows))) return rc;
  if ((rc = train_rows(n, n->pos, nullptr, n->targets, count, rows, lr, update, n->loss))) return rc;
  if (grads)
    MI_HIP(ctx, hipMemcpyAsync(grads, n->g, sizeof(float) * MI_MLP_PARAMS, hipMemcpyDeviceToHost,
                               ctx->stream));
  if (loss) MI_HIP(ctx, hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MI_OK;
}

int mi_mlp_gradients(mi_mlp *n, const mi_pos *pos, const float *targets, uint32_t count,
                     float *grads, float *loss) {
  if (!n || !grads) return MI_ERR_INVALID_ARGUMENT;
  return grad_common(n, pos, targets, count, 0.f, 0, grads, loss);
}

int mi_mlp_train_step(mi_mlp *n, const mi_pos *pos, const float *targets, uint32_t count,
                      float lr, float *loss) {
  if (!n) return MI_ERR_INVALID_ARGUMENT;
  return grad_common(n, pos, targets, count, lr, 1, nullptr, loss);
}

int mi_mlp_adam_state(mi_mlp *n, float *m, float *v, uint32_t *step) {
  if (!n) return MI_ERR_INVALID_ARGUMENT;
  mi_ctx *ctx = n->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (int rc = use_device(n)) return rc;
  if (m) MI_HIP(ctx, hipMemcpyAsync(m, n->m, sizeof(float) * MI_MLP_PARAMS, hipMemcpyDeviceToHost, ctx->stream));
  if (v) MI_HIP(ctx, hipMemcpyAsync(v, n->v, sizeof(float) * MI_MLP_PARAMS, hipMemcpyDeviceToHost, ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (step) *step = n->step;
  return MI_OK;
}

int mi_mlp_set_dataset(mi_mlp *n, const mi_pos *pos, const float *targets, uint32_t count) {
  if (!n || (count && (!pos || !targets))) return MI_ERR_INVALID_ARGUMENT;
  mi_ctx *ctx = n->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  int rc;
  if ((rc = use_device(n))) return rc;
  if ((rc = ensure(ctx, n->ds_pos, sizeof(mi_pos) * (size_t)std::max<uint32_t>(count, 1)))) return rc;
  if ((rc = ensure(ctx, n->ds_t, sizeof(float) * (size_t)std::max<uint32_t>(count, 1)))) return rc;
  if ((rc = ensure(ctx, n->ds_perm, sizeof(uint32_t) * (size_t)std::max<uint32_t>(count, 1)))) return rc;
  MI_HIP(ctx, hipMemcpyAsync(n->ds_pos.ptr, pos, sizeof(mi_pos) * count, hipMemcpyHostToDevice, ctx->stream));
  MI_HIP(ctx, hipMemcpyAsync(n->ds_t.ptr, targets, sizeof(float) * count, hipMemcpyHostToDevice, ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  n->ds_n = count;
  return MI_OK;
}

int mi_mlp_train_epoch(mi_mlp *n, uint32_t batch, float lr, uint32_t shuffle_seed, uint32_t epoch,
                       float *mean_loss, float *ms) {
  if (!n) return MI_ERR_INVALID_ARGUMENT;
  mi_ctx *ctx = n->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  if (batch == 0 || batch % 64 || batch > n->max_rows)
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "batch must be a multiple of 64 within max_batch");
  const uint32_t nb = n->ds_n / batch;
  if (nb == 0) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "the dataset holds less than one batch");
  if (int rc0 = use_device(n)) return rc0;
  // Fisher-Yates with Philox key (shuffle_seed, epoch), counter (i >> 2, 0, 3, 1)
  std::vector<uint32_t> perm(n->ds_n);
  std::iota(perm.begin(), perm.end(), 0u);
  uint32_t words[4];
  for (uint32_t i = n->ds_n - 1, c = 0; i > 0; i--, c++) {
    if ((c & 3) == 0) ck::philox(c >> 2, 0u, ck::TAG_MLP, 1u, shuffle_seed, epoch, words);
    std::swap(perm[i], perm[words[c & 3] % (i + 1)]);
  }
  int rc;
  if ((rc = ensure(ctx, n->ds_loss, sizeof(float) * nb))) return rc;
  MI_HIP(ctx, hipMemcpyAsync(n->ds_perm.ptr, perm.data(), sizeof(uint32_t) * n->ds_n,
                             hipMemcpyHostToDevice, ctx->stream));
  float *losses = (float *)n->ds_loss.ptr;
  MI_HIP(ctx, hipEventRecord(n->e0, ctx->stream));
  for (uint32_t b = 0; b < nb; b++) {
    const uint32_t *idx = (const uint32_t *)n->ds_perm.ptr + (size_t)b * batch;
    if ((rc = train_rows(n, (const mi_pos *)n->ds_pos.ptr, idx, (const float *)n->ds_t.ptr, batch,
                         batch, lr, 1, losses + b)))
      return rc;
  }
  MI_HIP(ctx, hipEventRecord(n->e1, ctx->stream));
  std::vector<float> l(nb);
  MI_HIP(ctx, hipMemcpyAsync(l.data(), losses, sizeof(float) * nb, hipMemcpyDeviceToHost, ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (mean_loss) *mean_loss = std::accumulate(l.begin(), l.end(), 0.f) / (float)nb;
  if (ms) MI_HIP(ctx, hipEventElapsedTime(ms, n->e0, n->e1));
  return MI_OK;
}

int mi_mlp_save(mi_mlp *n, const char *path) {
  if (!n || !path) return MI_ERR_INVALID_ARGUMENT;
  std::vector<float> w(MI_MLP_PARAMS), m(MI_MLP_PARAMS), v(MI_MLP_PARAMS);
  uint32_t step = 0;
  int rc;
  if ((rc = mi_mlp_get_params(n, w.data()))) return rc;
  if ((rc = mi_mlp_adam_state(n, m.data(), v.data(), &step))) return rc;
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(n->ctx, MI_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
  const char magic[8] = {'G', 'P', 'U', 'A', 'I', 'M', 'L', 'P'};
  const uint32_t head[3] = {1u, step, (uint32_t)MI_MLP_PARAMS};
  bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(head, 4, 3, f) == 3 &&
            std::fwrite(w.data(), 4, MI_MLP_PARAMS, f) == MI_MLP_PARAMS &&
            std::fwrite(m.data(), 4, MI_MLP_PARAMS, f) == MI_MLP_PARAMS &&
            std::fwrite(v.data(), 4, MI_MLP_PARAMS, f) == MI_MLP_PARAMS;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? MI_OK : fail(n->ctx, MI_ERR_INVALID_ARGUMENT, std::string("short write to ") + path);
}

int mi_mlp_load(mi_mlp *n, const char *path) {
  if (!n || !path) return MI_ERR_INVALID_ARGUMENT;
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(n->ctx, MI_ERR_INVALID_ARGUMENT, std::string("cannot read ") + path);
  char magic[8];
  uint32_t head[3];
  std::vector<float> w(MI_MLP_PARAMS), m(MI_MLP_PARAMS), v(MI_MLP_PARAMS);
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "GPUAIMLP", 8) == 0 &&
            std::fread(head, 4, 3, f) == 3 && head[0] == 1u && head[2] == MI_MLP_PARAMS &&
            std::fread(w.data(), 4, MI_MLP_PARAMS, f) == MI_MLP_PARAMS &&
            std::fread(m.data(), 4, MI_MLP_PARAMS, f) == MI_MLP_PARAMS &&
            std::fread(v.data(), 4, MI_MLP_PARAMS, f) == MI_MLP_PARAMS;
  std::fclose(f);
  if (!ok) return fail(n->ctx, MI_ERR_INVALID_ARGUMENT, std::string("not a checkpoint: ") + path);
  mi_ctx *ctx = n->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  int rc = use_device(n);
  if (!rc) rc = upload_params(n, w.data());
  if (rc) return rc;
  MI_HIP(ctx, hipMemcpyAsync(n->m, m.data(), sizeof(float) * MI_MLP_PARAMS, hipMemcpyHostToDevice, ctx->stream));
  MI_HIP(ctx, hipMemcpyAsync(n->v, v.data(), sizeof(float) * MI_MLP_PARAMS, hipMemcpyHostToDevice, ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  n->step = head[1];
  return MI_OK;
}

}  // extern "C"
