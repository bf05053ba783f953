// ctx.h -- the context object behind mi_ctx (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "gpuai/mi_playout.h"

namespace gpuai {
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
};
// Pinned host memory (hipHostMalloc): copies from and to it run asynchronously on the copy
// engines, where copies of pageable memory are staged through the runtime's own buffers.
struct HostBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
};
}  // namespace gpuai

// Default launch shape: 64-thread blocks, 12 per CU (3 waves per SIMD), queue refill, defer_min
// 16.  Among the fastest shapes for a 1M-playout batch, and a small batch spreads over many CUs
// instead of sharing a few SIMDs.
constexpr mi_tuning kDefaultTuning = {64u, 12u, MI_REFILL_QUEUE, 16u};

// Launch-shape table: the fastest shape the autotuner measured for a policy, a batch-size bucket
// (floor(log2(playouts))) and a refill policy; kernel_ms < 0 marks an empty entry.
constexpr uint32_t kTuneBuckets = 32;
struct TuneEntry {
  mi_tuning tuning = {0u, 0u, 0u, 0u};
  float kernel_ms = -1.f;
  uint32_t playouts = 0;  // size of the batch it was measured on
};

struct mi_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  mi_tuning tuning = kDefaultTuning;  // the shape of batches whose bucket the table lacks
  TuneEntry tune_table[3][kTuneBuckets][2];  // [policy][bucket][refill]
  uint32_t cus = 0, wave = 64, clock_khz = 0;
  gpuai::DevBuf pos, offs, winners, plies, hashes, work, tallies, moves, counts, keys;
  // longest-first order of the uploaded batch (see upload_batch): positions permuted into it,
  // the original index of each, and (uneven counts) prefix offsets in that order
  gpuai::DevBuf ord_pos, order, ord_offs;
  bool ordered = false;
  std::vector<uint32_t> host_order, host_ord_offs;
  std::vector<mi_pos> host_ord_pos;
  // pinned staging of mi_stage and of batches too large for the packed route: positions |
  // offsets on the way in, work words | winners on the way back.  Rewritten only while the
  // stream is idle.
  gpuai::HostBuf h_in, h_back;
  // the packed batch of mi_run_playouts (inputs | work words | winners in one device block and
  // its pinned mirror, one copy each way); used and released only by that route, which waits
  // for its stream before it returns
  gpuai::DevBuf batch;
  gpuai::HostBuf h_batch;
  // staged batch
  uint32_t staged_npos = 0, staged_per_pos = 0, staged_total = 0;
  bool staged = false, staged_offs = false;
  std::vector<uint32_t> host_offs;
  std::string last_error;
  std::mutex mu;
};

namespace gpuai {
namespace ctxutil {
// Records `msg` as the context's last error and returns `code`.
int fail(mi_ctx *ctx, int code, const std::string &msg);
// Records a HIP failure and returns MI_ERR_DEVICE or MI_ERR_OUT_OF_MEMORY.
int hip_fail(mi_ctx *ctx, hipError_t e, const char *what);
// Grows a device buffer to at least `bytes`.
int ensure(mi_ctx *ctx, DevBuf &b, size_t bytes);
// Grows a pinned host buffer to at least `bytes` (the caller's stream must be idle).
int ensure_host(mi_ctx *ctx, HostBuf &b, size_t bytes);
}  // namespace ctxutil
// Looks up the device address of the allocation-free entry point's scratch array for the
// context's device (once per device; the device must be current).
int noalloc_setup(mi_ctx *ctx);
}  // namespace gpuai

#define MI_HIP(ctx, call)                                                 \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return gpuai::ctxutil::hip_fail(ctx, e_, #call); \
  } while (0)
