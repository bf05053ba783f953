// capi.cpp -- the C ABI of include/gpuai/mi_playout.h.
//
// A context owns one HIP stream, two timing events and device buffers that grow on demand and
// are reused across calls (the reference allocates, copies 776 bytes per state and frees on
// every call, src/singlePlayout.cu:76-120).  Calls on one context are serialised by a mutex;
// separate contexts are independent, so each MCTS worker thread can own one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gpuai/mi_playout.h"
#include "kernels.h"

#include "ctx.h"

namespace gpuai {
namespace ctxutil {

int fail(mi_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

int hip_fail(mi_ctx *ctx, hipError_t e, const char *what) {
  std::string msg = std::string(what) + ": " + hipGetErrorString(e);
  int code = (e == hipErrorOutOfMemory) ? MI_ERR_OUT_OF_MEMORY : MI_ERR_DEVICE;
  return fail(ctx, code, msg);
}


int ensure(mi_ctx *ctx, DevBuf &b, size_t bytes) {
  if (bytes <= b.bytes && b.ptr) return MI_OK;
  if (b.ptr) {
    hipError_t e = hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return hip_fail(ctx, e, "hipFree");
  }
  size_t want = std::max<size_t>(bytes, 256);
  want = (want + 4095) & ~(size_t)4095;
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
  b.bytes = want;
  return MI_OK;
}

int ensure_host(mi_ctx *ctx, HostBuf &b, size_t bytes) {
  if (bytes <= b.bytes && b.ptr) return MI_OK;
  if (b.ptr) {
    hipError_t e = hipHostFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return hip_fail(ctx, e, "hipHostFree");
  }
  size_t want = std::max<size_t>(bytes, 4096);
  want = (want + 4095) & ~(size_t)4095;
  hipError_t e = hipHostMalloc(&b.ptr, want, hipHostMallocDefault);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipHostMalloc");
  b.bytes = want;
  return MI_OK;
}

}  // namespace ctxutil
}  // namespace gpuai

using gpuai::DevBuf;
using gpuai::ctxutil::ensure;
using gpuai::ctxutil::ensure_host;
using gpuai::ctxutil::fail;
using gpuai::ctxutil::hip_fail;

namespace {

int check_tuning(mi_ctx *ctx, const mi_tuning &t) {
  if (t.block_size < 64 || t.block_size > 1024 || t.block_size % 64)
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "block_size must be a multiple of 64 in [64, 1024]");
  if (t.blocks_per_cu < 1 || t.blocks_per_cu > 64)
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "blocks_per_cu must be in [1, 64]");
  if (t.refill != MI_REFILL_STATIC && t.refill != MI_REFILL_QUEUE)
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "unknown refill policy");
  if (t.defer_min < 1 || t.defer_min > 64)
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "defer_min must be in [1, 64]");
  return MI_OK;
}

// Validates a batch and computes playout offsets; total must stay below 2^31.
int plan_batch(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
               uint32_t per_pos, uint64_t &total) {
  if (npos > 0 && !pos) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "positions pointer is null");
  total = 0;
  if (counts) {
    ctx->host_offs.resize((size_t)npos + 1);
    ctx->host_offs[0] = 0;
    for (uint32_t i = 0; i < npos; i++) {
      total += counts[i];
      if (total >= (1ull << 31)) break;
      ctx->host_offs[i + 1] = (uint32_t)total;
    }
  } else {
    total = (uint64_t)npos * per_pos;
  }
  if (total >= (1ull << 31))
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "more than 2^31-1 playouts in one call");
  for (uint32_t i = 0; i < npos; i++) {
    const mi_pos &q = pos[i];
    if ((q.p1 & q.p2) || q.turn > 1)
      return fail(ctx, MI_ERR_INVALID_ARGUMENT,
                  "position " + std::to_string(i) + " is malformed (overlapping sides or turn > 1)");
  }
  return MI_OK;
}

// Longest-first hand-out.  At the end of a batch the chip idles while the last long games
// finish, one ply at a time per wave; the more pieces a position has, the longer its games
// tend to run, so positions are handed out by decreasing piece count (stable, counting sort)
// for batches of at least 256 playouts per CU (measured: 6-9% off a 262k-playout launch, no
// change at 65k, where nearly every playout starts at once).  Only the schedule changes: every
// playout keeps its stream and its result slot.
bool wants_order(const mi_ctx *ctx, uint32_t npos, uint64_t total) {
  return npos > 1 && total >= (uint64_t)ctx->cus * 256u;
}

void plan_order(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                uint64_t total) {
  ctx->ordered = wants_order(ctx, npos, total);
  if (!ctx->ordered) return;
  uint32_t start[34] = {0};  // bucket b holds positions with 32 - b pieces (b = 0: most)
  for (uint32_t i = 0; i < npos; i++) start[1 + (32u - (uint32_t)__builtin_popcount(pos[i].p1 | pos[i].p2))]++;
  for (uint32_t b = 1; b < 34; b++) start[b] += start[b - 1];
  ctx->host_order.resize(npos);
  ctx->host_ord_pos.resize(npos);
  for (uint32_t i = 0; i < npos; i++) {
    const uint32_t r = start[32u - (uint32_t)__builtin_popcount(pos[i].p1 | pos[i].p2)]++;
    ctx->host_order[r] = i;
    ctx->host_ord_pos[r] = pos[i];
  }
  if (counts) {
    ctx->host_ord_offs.resize((size_t)npos + 1);
    ctx->host_ord_offs[0] = 0;
    for (uint32_t r = 0; r < npos; r++)
      ctx->host_ord_offs[r + 1] = ctx->host_ord_offs[r] + counts[ctx->host_order[r]];
  }
}

// Pinned staging is used up to this many bytes per direction; larger batches copy straight from
// and to the caller's memory (their copies take far longer than the staging saves).
constexpr size_t kPinnedMax = size_t(64) << 20;

// Uploads a batch to the context's device buffers.  With `pinned` (and at most kPinnedMax
// bytes), positions and offsets go through the context's pinned staging (the stream is
// synchronised first, so that no earlier copy still reads it); otherwise straight from the
// caller's memory.
int upload_batch(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                 uint64_t total, bool pinned = false, const uint32_t *keys = nullptr) {
  int rc;
  plan_order(ctx, pos, npos, counts, total);
  if (ctx->ordered) {
    if ((rc = ensure(ctx, ctx->ord_pos, sizeof(mi_pos) * (size_t)npos))) return rc;
    if ((rc = ensure(ctx, ctx->order, sizeof(uint32_t) * (size_t)npos))) return rc;
    MI_HIP(ctx, hipMemcpyAsync(ctx->ord_pos.ptr, ctx->host_ord_pos.data(), sizeof(mi_pos) * npos,
                               hipMemcpyHostToDevice, ctx->stream));
    MI_HIP(ctx, hipMemcpyAsync(ctx->order.ptr, ctx->host_order.data(), sizeof(uint32_t) * npos,
                               hipMemcpyHostToDevice, ctx->stream));
    if (counts) {
      if ((rc = ensure(ctx, ctx->ord_offs, sizeof(uint32_t) * ((size_t)npos + 1)))) return rc;
      MI_HIP(ctx, hipMemcpyAsync(ctx->ord_offs.ptr, ctx->host_ord_offs.data(),
                                 sizeof(uint32_t) * ((size_t)npos + 1), hipMemcpyHostToDevice,
                                 ctx->stream));
    }
  }
  if ((rc = ensure(ctx, ctx->pos, sizeof(mi_pos) * (size_t)std::max<uint32_t>(npos, 1))))
    return rc;
  if ((rc = ensure(ctx, ctx->winners, (size_t)total))) return rc;
  if ((rc = ensure(ctx, ctx->work, sizeof(gpuai::WorkWords)))) return rc;
  if (counts && (rc = ensure(ctx, ctx->offs, sizeof(uint32_t) * ((size_t)npos + 1)))) return rc;
  if (keys && (rc = ensure(ctx, ctx->keys, sizeof(uint32_t) * (size_t)std::max<uint32_t>(npos, 1))))
    return rc;
  const size_t pos_bytes = sizeof(mi_pos) * npos;
  const size_t offs_bytes = counts ? sizeof(uint32_t) * ((size_t)npos + 1) : 0;
  const size_t keys_bytes = keys ? sizeof(uint32_t) * (size_t)npos : 0;
  const void *src_pos = pos, *src_offs = ctx->host_offs.data(), *src_keys = keys;
  if (pinned && pos_bytes + offs_bytes + keys_bytes <= kPinnedMax) {
    MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((rc = ensure_host(ctx, ctx->h_in, pos_bytes + offs_bytes + keys_bytes))) return rc;
    char *h = (char *)ctx->h_in.ptr;
    std::memcpy(h, pos, pos_bytes);
    if (counts) std::memcpy(h + pos_bytes, ctx->host_offs.data(), offs_bytes);
    if (keys) std::memcpy(h + pos_bytes + offs_bytes, keys, keys_bytes);
    src_pos = h;
    src_offs = h + pos_bytes;
    src_keys = h + pos_bytes + offs_bytes;
  }
  MI_HIP(ctx, hipMemcpyAsync(ctx->pos.ptr, src_pos, pos_bytes, hipMemcpyHostToDevice, ctx->stream));
  if (counts)
    MI_HIP(ctx, hipMemcpyAsync(ctx->offs.ptr, src_offs, offs_bytes, hipMemcpyHostToDevice,
                               ctx->stream));
  if (keys)
    MI_HIP(ctx, hipMemcpyAsync(ctx->keys.ptr, src_keys, keys_bytes, hipMemcpyHostToDevice,
                               ctx->stream));
  return MI_OK;
}

uint32_t grid_for(const mi_ctx *ctx, const mi_tuning &t, uint64_t total) {
  const uint64_t block = t.block_size;
  const uint64_t need = (total + block - 1) / block;
  const uint64_t resident = (uint64_t)ctx->cus * t.blocks_per_cu;
  return gpuai::grid_size(need, resident);
}

uint32_t tune_bucket(uint64_t total) {
  const uint32_t b = total ? 63u - (uint32_t)__builtin_clzll(total) : 0u;
  return b < kTuneBuckets ? b : kTuneBuckets - 1u;
}

// The launch shape of a batch of `total` playouts: the table's entry for its policy and bucket
// (the faster refill, or refill `only` when it is 0 or 1), else the context's tuning.
mi_tuning shape_for(const mi_ctx *ctx, uint32_t policy, uint64_t total, int only = -1) {
  const TuneEntry *e = ctx->tune_table[policy][tune_bucket(total)];
  const TuneEntry *best = nullptr;
  for (int r = 0; r < 2; r++) {
    if (e[r].kernel_ms < 0.f || (only >= 0 && r != only)) continue;
    if (!best || e[r].kernel_ms < best->kernel_ms) best = &e[r];
  }
  return best ? best->tuning : ctx->tuning;
}

void clear_tune_table(mi_ctx *ctx) {
  for (auto &pol : ctx->tune_table)
    for (auto &bucket : pol)
      for (TuneEntry &e : bucket) e = TuneEntry();
}

// Launches the playout kernel on the staged device buffers and reads back the statistics and,
// when `winners` is set, the winners (both through the context's pinned read-back buffer, with
// one synchronisation).
int run_kernel(mi_ctx *ctx, uint32_t npos, uint32_t per_pos, uint64_t total, bool with_offs,
               uint32_t seed, uint32_t epoch, uint32_t policy, bool trace, mi_stats *stats,
               int8_t *winners = nullptr, bool keyed = false, const mi_tuning *shape = nullptr) {
  // the autotuner passes the candidate it measures; every other launch takes its batch's shape
  const mi_tuning t = shape ? *shape : shape_for(ctx, policy, total);
  int rc;
  if ((rc = ensure_host(ctx, ctx->h_back, sizeof(gpuai::WorkWords) + (winners ? total : 0))))
    return rc;
  gpuai::PlayoutArgs a;
  a.pos = (const mi_pos *)ctx->pos.ptr;
  a.offs = with_offs ? (const uint32_t *)ctx->offs.ptr : nullptr;
  a.npos = npos;
  a.per_pos = per_pos;
  a.per_pos_shift = per_pos && !(per_pos & (per_pos - 1u)) ? (uint32_t)__builtin_ctz(per_pos) : 32u;
  a.total = (uint32_t)total;
  a.seed = seed;
  a.epoch = epoch;
  a.winners = (int8_t *)ctx->winners.ptr;
  a.plies = trace ? (uint32_t *)ctx->plies.ptr : nullptr;
  a.hashes = trace ? (uint32_t *)ctx->hashes.ptr : nullptr;
  a.work = (gpuai::WorkWords *)ctx->work.ptr;
  a.defer_min = t.defer_min;
  a.desc = nullptr;
  a.ord_pos = ctx->ordered ? (const mi_pos *)ctx->ord_pos.ptr : nullptr;
  a.order = ctx->ordered ? (const uint32_t *)ctx->order.ptr : nullptr;
  a.ord_offs = ctx->ordered && with_offs ? (const uint32_t *)ctx->ord_offs.ptr : nullptr;
  a.keys = keyed ? (const uint32_t *)ctx->keys.ptr : nullptr;
  const uint32_t grid = grid_for(ctx, t, total);
  MI_HIP(ctx, gpuai::launch_reset(a.work, ctx->stream));
  MI_HIP(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  MI_HIP(ctx, gpuai::launch_playouts(a, policy, trace, t.refill, grid, t.block_size, ctx->stream));
  MI_HIP(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
  gpuai::WorkWords *hw = (gpuai::WorkWords *)ctx->h_back.ptr;
  int8_t *hwin = (int8_t *)ctx->h_back.ptr + sizeof(gpuai::WorkWords);
  MI_HIP(ctx, hipMemcpyAsync(hw, a.work, sizeof(*hw), hipMemcpyDeviceToHost, ctx->stream));
  if (winners)
    MI_HIP(ctx, hipMemcpyAsync(hwin, a.winners, total, hipMemcpyDeviceToHost, ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const gpuai::WorkWords w = *hw;
  if (winners) std::memcpy(winners, hwin, total);
  float ms = 0.f;
  MI_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  if (stats) {
    stats->playouts = total;
    stats->plies = w.plies;
    stats->kernel_ms = ms;
    stats->max_moves = w.max_moves;
    stats->grid_blocks = grid;
    stats->block_size = t.block_size;
    stats->refill = t.refill;
  }
  return MI_OK;
}

int bad_policy(mi_ctx *ctx, uint32_t policy) {
  if (policy > MI_POLICY_HEURISTIC_NOISY) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "unknown policy");
  return MI_OK;
}

using Clock = std::chrono::steady_clock;

void finish_stats(mi_stats *stats, Clock::time_point t0) {
  if (stats)
    stats->total_ms = std::chrono::duration<float, std::milli>(Clock::now() - t0).count();
}

// The packed route of mi_run_playouts: one upload and one read-back per batch.  The context's
// batch buffer holds, in device memory and in its pinned mirror, the sections
//   positions | offsets | keys | longest-first positions | order | ordered offsets |
//   work words | winners
// (each 256-byte aligned; a section the batch does not use is empty), so that one copy moves the
// inputs together with zeroed work words, which resets the claim counters without a launch, and
// one copy brings the work words back with the winners.  Host-side time above the kernel per
// call, against three uploads (two staged by the runtime from pageable memory), a reset launch
// and two read-backs: 112 -> 85 us for 1M playouts, 59 -> 31 us for 65,536, 36 -> 27 us for
// 4095 over 2048 positions with uneven counts.
struct PackedLayout {
  size_t pos = 0, offs = 0, keys = 0, ord_pos = 0, order = 0, ord_offs = 0, work = 0, win = 0,
         end = 0;
};

PackedLayout packed_layout(uint32_t npos, bool with_offs, bool keyed, bool ordered,
                           uint64_t total) {
  PackedLayout L;
  size_t at = 0;
  auto take = [&at](size_t bytes) {
    const size_t r = at;
    at = (at + bytes + 255) & ~(size_t)255;
    return r;
  };
  const size_t n = npos;
  L.pos = take(ordered ? 0 : sizeof(mi_pos) * n);  // an ordered batch is read in rank order only
  L.offs = take(with_offs ? sizeof(uint32_t) * (n + 1) : 0);
  L.keys = take(keyed ? sizeof(uint32_t) * n : 0);
  L.ord_pos = take(ordered ? sizeof(mi_pos) * n : 0);
  L.order = take(ordered ? sizeof(uint32_t) * n : 0);
  L.ord_offs = take(ordered && with_offs ? sizeof(uint32_t) * (n + 1) : 0);
  L.work = take(sizeof(gpuai::WorkWords));
  L.win = take((size_t)total);
  L.end = at;
  return L;
}

// Runs a planned batch (plan_batch and plan_order done, total > 0) whose inputs and work words
// fit the pinned staging limit.  The winners come back through the mirror when they fit it too,
// else straight to the caller's memory.
int run_packed(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
               uint32_t per_pos, uint64_t total, const PackedLayout &L, uint32_t seed,
               uint32_t epoch, uint32_t policy, bool trace, int8_t *winners, uint32_t *plies,
               uint32_t *hashes, uint32_t *tallies, mi_stats *stats, const uint32_t *keys) {
  int rc;
  const bool back = winners && total <= kPinnedMax;
  if ((rc = ensure(ctx, ctx->batch, L.end))) return rc;
  if ((rc = ensure_host(ctx, ctx->h_batch, back ? L.end : L.win))) return rc;
  if (trace) {
    if ((rc = ensure(ctx, ctx->plies, sizeof(uint32_t) * total))) return rc;
    if ((rc = ensure(ctx, ctx->hashes, sizeof(uint32_t) * total))) return rc;
  }
  if (tallies && (rc = ensure(ctx, ctx->tallies, sizeof(uint32_t) * 3 * (size_t)npos))) return rc;
  char *h = (char *)ctx->h_batch.ptr;
  char *d = (char *)ctx->batch.ptr;
  const size_t n = npos;
  if (ctx->ordered) {
    std::memcpy(h + L.ord_pos, ctx->host_ord_pos.data(), sizeof(mi_pos) * n);
    std::memcpy(h + L.order, ctx->host_order.data(), sizeof(uint32_t) * n);
    if (counts) std::memcpy(h + L.ord_offs, ctx->host_ord_offs.data(), sizeof(uint32_t) * (n + 1));
  } else {
    std::memcpy(h + L.pos, pos, sizeof(mi_pos) * n);
  }
  if (counts) std::memcpy(h + L.offs, ctx->host_offs.data(), sizeof(uint32_t) * (n + 1));
  if (keys) std::memcpy(h + L.keys, keys, sizeof(uint32_t) * n);
  std::memset(h + L.work, 0, sizeof(gpuai::WorkWords));

  const mi_tuning t = shape_for(ctx, policy, total);
  gpuai::PlayoutArgs a;
  a.pos = ctx->ordered ? nullptr : (const mi_pos *)(d + L.pos);
  a.offs = counts ? (const uint32_t *)(d + L.offs) : nullptr;
  a.ord_pos = ctx->ordered ? (const mi_pos *)(d + L.ord_pos) : nullptr;
  a.order = ctx->ordered ? (const uint32_t *)(d + L.order) : nullptr;
  a.ord_offs = ctx->ordered && counts ? (const uint32_t *)(d + L.ord_offs) : nullptr;
  a.keys = keys ? (const uint32_t *)(d + L.keys) : nullptr;
  a.npos = npos;
  a.per_pos = per_pos;
  a.per_pos_shift = per_pos && !(per_pos & (per_pos - 1u)) ? (uint32_t)__builtin_ctz(per_pos) : 32u;
  a.total = (uint32_t)total;
  a.seed = seed;
  a.epoch = epoch;
  a.winners = (int8_t *)(d + L.win);
  a.plies = trace ? (uint32_t *)ctx->plies.ptr : nullptr;
  a.hashes = trace ? (uint32_t *)ctx->hashes.ptr : nullptr;
  a.work = (gpuai::WorkWords *)(d + L.work);
  a.defer_min = t.defer_min;
  a.desc = nullptr;
  const uint32_t grid = grid_for(ctx, t, total);
  MI_HIP(ctx, hipMemcpyAsync(d, h, L.win, hipMemcpyHostToDevice, ctx->stream));
  MI_HIP(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
  MI_HIP(ctx, gpuai::launch_playouts(a, policy, trace, t.refill, grid, t.block_size, ctx->stream));
  MI_HIP(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
  MI_HIP(ctx, hipMemcpyAsync(h + L.work, d + L.work, (back ? L.win + total : L.win) - L.work,
                             hipMemcpyDeviceToHost, ctx->stream));
  if (winners && !back)
    MI_HIP(ctx, hipMemcpyAsync(winners, d + L.win, total, hipMemcpyDeviceToHost, ctx->stream));
  if (trace && plies)
    MI_HIP(ctx, hipMemcpyAsync(plies, ctx->plies.ptr, sizeof(uint32_t) * total,
                               hipMemcpyDeviceToHost, ctx->stream));
  if (trace && hashes)
    MI_HIP(ctx, hipMemcpyAsync(hashes, ctx->hashes.ptr, sizeof(uint32_t) * total,
                               hipMemcpyDeviceToHost, ctx->stream));
  if (tallies) {
    MI_HIP(ctx, gpuai::launch_tally(a.winners, a.offs, npos, per_pos,
                                    (uint32_t *)ctx->tallies.ptr, ctx->stream));
    MI_HIP(ctx, hipMemcpyAsync(tallies, ctx->tallies.ptr, sizeof(uint32_t) * 3 * n,
                               hipMemcpyDeviceToHost, ctx->stream));
  }
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const gpuai::WorkWords *w = (const gpuai::WorkWords *)(h + L.work);
  if (back) std::memcpy(winners, h + L.win, total);
  float ms = 0.f;
  MI_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  if (stats) {
    stats->playouts = total;
    stats->plies = w->plies;
    stats->kernel_ms = ms;
    stats->max_moves = w->max_moves;
    stats->grid_blocks = grid;
    stats->block_size = t.block_size;
    stats->refill = t.refill;
  }
  return MI_OK;
}

int run_common(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
               uint32_t per_pos, uint32_t seed, uint32_t epoch, uint32_t policy, bool trace,
               int8_t *winners, uint32_t *plies, uint32_t *hashes, uint32_t *tallies,
               mi_stats *stats, const uint32_t *keys = nullptr) {
  if (!ctx) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  const auto t0 = Clock::now();
  ctx->last_error.clear();
  if (stats) std::memset(stats, 0, sizeof(*stats));
  int rc;
  if ((rc = bad_policy(ctx, policy))) return rc;
  uint64_t total = 0;
  if ((rc = plan_batch(ctx, pos, npos, counts, per_pos, total))) return rc;
  if (tallies && npos) std::memset(tallies, 0, sizeof(uint32_t) * 3 * (size_t)npos);
  if (total == 0) {  // src/singlePlayout.cu:73-75: an empty batch returns at once
    finish_stats(stats, t0);
    return MI_OK;
  }
  if (!winners && !tallies) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "winners pointer is null");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  const PackedLayout L = packed_layout(npos, counts != nullptr, keys != nullptr,
                                       wants_order(ctx, npos, total), total);
  if (L.win <= kPinnedMax) {
    plan_order(ctx, pos, npos, counts, total);
    ctx->staged = false;  // the order plan now describes this call's batch
    if ((rc = run_packed(ctx, pos, npos, counts, per_pos, total, L, seed, epoch, policy, trace,
                         winners, plies, hashes, tallies, stats, keys)))
      return rc;
    finish_stats(stats, t0);
    return MI_OK;
  }
  // inputs beyond the pinned staging limit: each section straight from the caller's memory
  if ((rc = upload_batch(ctx, pos, npos, counts, total, true, keys))) return rc;
  if (trace) {
    if ((rc = ensure(ctx, ctx->plies, sizeof(uint32_t) * total))) return rc;
    if ((rc = ensure(ctx, ctx->hashes, sizeof(uint32_t) * total))) return rc;
  }
  ctx->staged = false;  // the batch buffers now hold this call's data
  const bool pinned_back = total <= kPinnedMax;
  if ((rc = run_kernel(ctx, npos, per_pos, total, counts != nullptr, seed, epoch, policy, trace,
                       stats, pinned_back ? winners : nullptr, keys != nullptr)))
    return rc;
  if (winners && !pinned_back)
    MI_HIP(ctx, hipMemcpyAsync(winners, ctx->winners.ptr, total, hipMemcpyDeviceToHost,
                               ctx->stream));
  if (trace && plies)
    MI_HIP(ctx, hipMemcpyAsync(plies, ctx->plies.ptr, sizeof(uint32_t) * total,
                               hipMemcpyDeviceToHost, ctx->stream));
  if (trace && hashes)
    MI_HIP(ctx, hipMemcpyAsync(hashes, ctx->hashes.ptr, sizeof(uint32_t) * total,
                               hipMemcpyDeviceToHost, ctx->stream));
  if (tallies) {
    if ((rc = ensure(ctx, ctx->tallies, sizeof(uint32_t) * 3 * (size_t)npos))) return rc;
    MI_HIP(ctx, gpuai::launch_tally((const int8_t *)ctx->winners.ptr,
                                    counts ? (const uint32_t *)ctx->offs.ptr : nullptr, npos,
                                    per_pos, (uint32_t *)ctx->tallies.ptr, ctx->stream));
    MI_HIP(ctx, hipMemcpyAsync(tallies, ctx->tallies.ptr, sizeof(uint32_t) * 3 * (size_t)npos,
                               hipMemcpyDeviceToHost, ctx->stream));
  }
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  finish_stats(stats, t0);
  return MI_OK;
}

}  // namespace

extern "C" {

const char *mi_strerror(int code) {
  switch (code) {
    case MI_OK: return "ok";
    case MI_ERR_INVALID_ARGUMENT: return "invalid argument";
    case MI_ERR_NO_DEVICE: return "no HIP device";
    case MI_ERR_DEVICE: return "device error";
    case MI_ERR_OUT_OF_MEMORY: return "out of device memory";
    case MI_ERR_CAPACITY: return "move list longer than the capacity";
    case MI_ERR_NOT_STAGED: return "no staged batch";
    default: return "unknown error";
  }
}

const char *mi_last_error(const mi_ctx *ctx) { return ctx ? ctx->last_error.c_str() : ""; }

int mi_create(mi_ctx **out, int device, const mi_tuning *tuning) {
  if (!out) return MI_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MI_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return MI_ERR_INVALID_ARGUMENT;
  mi_ctx *ctx = new mi_ctx();
  ctx->device = device;
  if (tuning) {
    mi_tuning t = *tuning;
    if (!t.block_size) t.block_size = ctx->tuning.block_size;
    if (!t.blocks_per_cu) t.blocks_per_cu = ctx->tuning.blocks_per_cu;
    if (!t.defer_min) t.defer_min = ctx->tuning.defer_min;
    if (check_tuning(ctx, t)) {
      delete ctx;
      return MI_ERR_INVALID_ARGUMENT;
    }
    ctx->tuning = t;
  }
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev_start) != hipSuccess || hipEventCreate(&ctx->ev_stop) != hipSuccess) {
    mi_destroy(ctx);
    return MI_ERR_DEVICE;
  }
  ctx->cus = (uint32_t)prop.multiProcessorCount;
  ctx->wave = (uint32_t)prop.warpSize;
  ctx->clock_khz = (uint32_t)prop.clockRate;
  if (gpuai::noalloc_setup(ctx) != MI_OK) {
    mi_destroy(ctx);
    return MI_ERR_DEVICE;
  }
  *out = ctx;
  return MI_OK;
}

int mi_destroy(mi_ctx *ctx) {
  if (!ctx) return MI_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DevBuf *bufs[] = {&ctx->pos,   &ctx->offs,    &ctx->winners, &ctx->plies,  &ctx->hashes,
                    &ctx->work,  &ctx->tallies, &ctx->moves,   &ctx->counts, &ctx->ord_pos,
                    &ctx->order, &ctx->ord_offs, &ctx->keys,   &ctx->batch};
  for (DevBuf *b : bufs)
    if (b->ptr) (void)hipFree(b->ptr);
  if (ctx->h_in.ptr) (void)hipHostFree(ctx->h_in.ptr);
  if (ctx->h_back.ptr) (void)hipHostFree(ctx->h_back.ptr);
  if (ctx->h_batch.ptr) (void)hipHostFree(ctx->h_batch.ptr);
  if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
  if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MI_OK;
}

int mi_set_tuning(mi_ctx *ctx, const mi_tuning *tuning) {
  if (!ctx || !tuning) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  mi_tuning t = *tuning;
  if (!t.block_size) t.block_size = kDefaultTuning.block_size;
  if (!t.blocks_per_cu) t.blocks_per_cu = kDefaultTuning.blocks_per_cu;
  if (!t.defer_min) t.defer_min = kDefaultTuning.defer_min;
  int rc = check_tuning(ctx, t);
  if (rc) return rc;
  ctx->tuning = t;
  clear_tune_table(ctx);  // an explicit shape holds for every batch size
  return MI_OK;
}

int mi_get_tuning(const mi_ctx *ctx, mi_tuning *tuning) {
  if (!ctx || !tuning) return MI_ERR_INVALID_ARGUMENT;
  *tuning = ctx->tuning;
  return MI_OK;
}

int mi_device_info(const mi_ctx *ctx, uint32_t *compute_units, uint32_t *wave_size,
                   uint32_t *clock_khz) {
  if (!ctx) return MI_ERR_INVALID_ARGUMENT;
  if (compute_units) *compute_units = ctx->cus;
  if (wave_size) *wave_size = ctx->wave;
  if (clock_khz) *clock_khz = ctx->clock_khz;
  return MI_OK;
}

int mi_run_playouts(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                    uint32_t per_pos, uint32_t seed, uint32_t epoch, uint32_t policy,
                    int8_t *winners, mi_stats *stats) {
  return run_common(ctx, pos, npos, counts, per_pos, seed, epoch, policy, false, winners, nullptr,
                    nullptr, nullptr, stats);
}

int mi_run_playouts_keyed(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                          uint32_t per_pos, const uint32_t *keys, uint32_t seed, uint32_t epoch,
                          uint32_t policy, int8_t *winners, mi_stats *stats) {
  return run_common(ctx, pos, npos, counts, per_pos, seed, epoch, policy, false, winners, nullptr,
                    nullptr, nullptr, stats, keys);
}

int mi_run_playouts_traced(mi_ctx *ctx, const mi_pos *pos, uint32_t npos,
                           const uint32_t *counts, uint32_t per_pos, uint32_t seed,
                           uint32_t epoch, uint32_t policy, int8_t *winners, uint32_t *plies,
                           uint32_t *hashes, mi_stats *stats) {
  return run_common(ctx, pos, npos, counts, per_pos, seed, epoch, policy, true, winners, plies,
                    hashes, nullptr, stats);
}

int mi_run_playouts_tally(mi_ctx *ctx, const mi_pos *pos, uint32_t npos,
                          const uint32_t *counts, uint32_t per_pos, uint32_t seed,
                          uint32_t epoch, uint32_t policy, uint32_t *tallies, mi_stats *stats) {
  if (!tallies && npos) return MI_ERR_INVALID_ARGUMENT;
  return run_common(ctx, pos, npos, counts, per_pos, seed, epoch, policy, false, nullptr, nullptr,
                    nullptr, tallies, stats);
}

int mi_gen_moves(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, mi_move *moves, uint32_t cap,
                 uint32_t *counts) {
  if (!ctx) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  if (npos == 0) return MI_OK;
  if (!pos || !moves || !counts || cap == 0)
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "null buffer or zero capacity");
  uint64_t dummy = 0;
  int rc;
  if ((rc = plan_batch(ctx, pos, npos, nullptr, 0, dummy))) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  const size_t move_bytes = sizeof(mi_move) * (size_t)npos * cap;
  if ((rc = ensure(ctx, ctx->pos, sizeof(mi_pos) * npos))) return rc;
  if ((rc = ensure(ctx, ctx->moves, move_bytes))) return rc;
  if ((rc = ensure(ctx, ctx->counts, sizeof(uint32_t) * npos))) return rc;
  if ((rc = ensure(ctx, ctx->work, sizeof(gpuai::WorkWords)))) return rc;
  ctx->staged = false;
  MI_HIP(ctx, hipMemcpyAsync(ctx->pos.ptr, pos, sizeof(mi_pos) * npos, hipMemcpyHostToDevice,
                             ctx->stream));
  gpuai::WorkWords *work = (gpuai::WorkWords *)ctx->work.ptr;
  MI_HIP(ctx, gpuai::launch_reset(work, ctx->stream));
  MI_HIP(ctx, gpuai::launch_gen_moves((const mi_pos *)ctx->pos.ptr, npos, (mi_move *)ctx->moves.ptr,
                                      cap, (uint32_t *)ctx->counts.ptr, work, ctx->stream));
  gpuai::WorkWords w;
  MI_HIP(ctx, hipMemcpyAsync(&w, work, sizeof(w), hipMemcpyDeviceToHost, ctx->stream));
  MI_HIP(ctx, hipMemcpyAsync(counts, ctx->counts.ptr, sizeof(uint32_t) * npos,
                             hipMemcpyDeviceToHost, ctx->stream));
  MI_HIP(ctx, hipMemcpyAsync(moves, ctx->moves.ptr, move_bytes, hipMemcpyDeviceToHost,
                             ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (w.overflow) return fail(ctx, MI_ERR_CAPACITY, "a move list is longer than cap");
  return MI_OK;
}

int mi_gen_start_positions(mi_ctx *ctx, uint32_t seed, uint32_t n, uint32_t kmin, uint32_t kmax,
                           uint32_t reject_terminal, mi_pos *out) {
  if (!ctx) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  if (n == 0) return MI_OK;
  if (!out || kmax < kmin) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "null output or kmax < kmin");
  if (kmax - kmin == 0xFFFFFFFFu)  // the count of ply numbers, kmax - kmin + 1, must fit 32 bits
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "kmax - kmin + 1 exceeds 2^32 - 1");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  int rc;
  if ((rc = ensure(ctx, ctx->pos, sizeof(mi_pos) * (size_t)n))) return rc;
  if ((rc = ensure(ctx, ctx->work, sizeof(gpuai::WorkWords)))) return rc;
  ctx->staged = false;
  gpuai::WorkWords *work = (gpuai::WorkWords *)ctx->work.ptr;
  MI_HIP(ctx, gpuai::launch_reset(work, ctx->stream));
  MI_HIP(ctx, gpuai::launch_start_positions(seed, n, kmin, kmax, reject_terminal,
                                            (mi_pos *)ctx->pos.ptr, work, ctx->stream));
  gpuai::WorkWords w;
  MI_HIP(ctx, hipMemcpyAsync(&w, work, sizeof(w), hipMemcpyDeviceToHost, ctx->stream));
  MI_HIP(ctx, hipMemcpyAsync(out, ctx->pos.ptr, sizeof(mi_pos) * (size_t)n, hipMemcpyDeviceToHost,
                             ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (w.overflow) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "no unfinished position in 4096 attempts");
  return MI_OK;
}

namespace {

// Measures every candidate launch shape on the uploaded batch -- workgroup sizes 64..1024, 1..8
// resident waves per SIMD, both refill policies, deferral thresholds 1, 8, 16 and 32 -- `reps`
// launches each (median kernel time), records the fastest of each refill policy in the tuning
// table's entry for the batch's policy and size bucket, and returns the fastest overall.
int tune_sweep(mi_ctx *ctx, uint32_t npos, uint32_t per_pos, uint64_t total, bool with_offs,
               uint32_t policy, uint32_t reps, mi_tune_result *results, uint32_t max_results,
               uint32_t *nresults, mi_tuning &best_t, float &best_ms) {
  const uint32_t blocks[] = {64u, 128u, 192u, 256u, 512u, 768u, 1024u};
  // 3 and 4 waves per SIMD are the fast ones on random-1024x1024 (more lanes in flight lengthen
  // the end of a batch, fewer leave the SIMDs short of waves to issue from)
  const uint32_t waves_per_simd[] = {1u, 2u, 3u, 4u, 8u};
  // 32 for the noisy heuristic, whose deferred plies make the draws: 1.07 against 1.16-1.19 ms
  // at 16 on noisy-1024x1024 (the other policies gain nothing past 16)
  const uint32_t defers[] = {1u, 8u, 16u, 32u};
  mi_tuning refill_best[2] = {};
  float refill_ms[2] = {3.0e38f, 3.0e38f};
  uint32_t n = 0;
  for (uint32_t refill = 0; refill < 2; refill++) {
    for (uint32_t b : blocks) {
      for (uint32_t w : waves_per_simd) {
        for (uint32_t dm : defers) {
          const uint32_t waves_per_cu = 4u * w;   // four SIMDs per CU
          if ((b / 64u) > waves_per_cu) continue;  // a block larger than the residency
          if ((waves_per_cu * 64u) % b) continue;  // whole blocks only
          const mi_tuning t = {b, waves_per_cu * 64u / b, refill, dm};
          std::vector<float> ms;
          for (uint32_t r = 0; r < reps; r++) {
            mi_stats st;
            const int rc = run_kernel(ctx, npos, per_pos, total, with_offs, 0x7A5Eu,
                                      0xFFFF0000u + r, policy, false, &st, nullptr, false, &t);
            if (rc) return rc;
            ms.push_back(st.kernel_ms);
          }
          std::sort(ms.begin(), ms.end());
          const float med = ms[ms.size() / 2];
          if (results && n < max_results) {
            results[n].tuning = t;
            results[n].kernel_ms = med;
            results[n].playouts_per_s = (float)((double)total / (med * 1e-3));
          }
          n++;
          if (med < refill_ms[refill]) {
            refill_ms[refill] = med;
            refill_best[refill] = t;
          }
        }
      }
    }
  }
  TuneEntry *e = ctx->tune_table[policy][tune_bucket(total)];
  for (int r = 0; r < 2; r++) {
    e[r].tuning = refill_best[r];
    e[r].kernel_ms = refill_ms[r];
    e[r].playouts = (uint32_t)total;
  }
  const int r = refill_ms[1] < refill_ms[0] ? 1 : 0;
  best_t = refill_best[r];
  best_ms = refill_ms[r];
  if (nresults) *nresults = n;
  return MI_OK;
}

}  // namespace

int mi_autotune(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                uint32_t per_pos, uint32_t policy, uint32_t reps, mi_tune_result *results,
                uint32_t max_results, uint32_t *nresults, mi_tuning *best) {
  if (!ctx) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  int rc;
  if ((rc = bad_policy(ctx, policy))) return rc;
  if (reps == 0) reps = 3;
  uint64_t total = 0;
  if ((rc = plan_batch(ctx, pos, npos, counts, per_pos, total))) return rc;
  if (total == 0) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "empty tuning batch");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  if ((rc = upload_batch(ctx, pos, npos, counts, total))) return rc;
  ctx->staged = false;
  mi_tuning best_t = ctx->tuning;
  float best_ms = 0.f;
  if ((rc = tune_sweep(ctx, npos, per_pos, total, counts != nullptr, policy, reps, results,
                       max_results, nresults, best_t, best_ms)))
    return rc;
  ctx->tuning = best_t;
  if (best) *best = best_t;
  return MI_OK;
}

int mi_autotune_sizes(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, uint32_t policy,
                      const uint32_t *sizes, uint32_t nsizes, uint32_t reps, mi_tune_entry *best) {
  if (!ctx) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  int rc;
  if ((rc = bad_policy(ctx, policy))) return rc;
  if (npos == 0 || !pos || (nsizes && !sizes))
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "no positions or sizes");
  if (reps == 0) reps = 3;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  for (uint32_t i = 0; i < nsizes; i++) {
    if (sizes[i] == 0) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "a batch size is 0");
    // the first positions, each with the same playout count, making up the size (rounded down)
    const uint32_t n = sizes[i] < npos ? sizes[i] : npos;
    const uint32_t per = sizes[i] / n;
    uint64_t total = 0;
    if ((rc = plan_batch(ctx, pos, n, nullptr, per, total))) return rc;
    if ((rc = upload_batch(ctx, pos, n, nullptr, total))) return rc;
    ctx->staged = false;
    mi_tuning t = ctx->tuning;
    float ms = 0.f;
    if ((rc = tune_sweep(ctx, n, per, total, false, policy, reps, nullptr, 0, nullptr, t, ms)))
      return rc;
    if (best) {
      best[i].policy = policy;
      best[i].bucket = tune_bucket(total);
      best[i].tuning = t;
      best[i].kernel_ms = ms;
      best[i].playouts = (uint32_t)total;
    }
  }
  return MI_OK;
}

int mi_get_tuning_table(const mi_ctx *ctx, mi_tune_entry *out, uint32_t cap, uint32_t *count) {
  if (!ctx || !count || (cap && !out)) return MI_ERR_INVALID_ARGUMENT;
  uint32_t n = 0;
  for (uint32_t pol = 0; pol <= MI_POLICY_HEURISTIC_NOISY; pol++)
    for (uint32_t b = 0; b < kTuneBuckets; b++)
      for (uint32_t r = 0; r < 2; r++) {
        const TuneEntry &e = ctx->tune_table[pol][b][r];
        if (e.kernel_ms < 0.f) continue;
        if (n < cap) {
          out[n].policy = pol;
          out[n].bucket = b;
          out[n].tuning = e.tuning;
          out[n].kernel_ms = e.kernel_ms;
          out[n].playouts = e.playouts;
        }
        n++;
      }
  *count = n;
  return MI_OK;
}

int mi_stage(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
             uint32_t per_pos) {
  if (!ctx) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  uint64_t total = 0;
  int rc;
  if ((rc = plan_batch(ctx, pos, npos, counts, per_pos, total))) return rc;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  if (total > 0 && (rc = upload_batch(ctx, pos, npos, counts, total))) return rc;
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->staged = true;
  ctx->staged_offs = counts != nullptr;
  ctx->staged_npos = npos;
  ctx->staged_per_pos = per_pos;
  ctx->staged_total = (uint32_t)total;
  return MI_OK;
}

int mi_launch_staged(mi_ctx *ctx, uint32_t seed, uint32_t epoch, uint32_t policy,
                     mi_stats *stats) {
  if (!ctx) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  const auto t0 = Clock::now();
  ctx->last_error.clear();
  if (stats) std::memset(stats, 0, sizeof(*stats));
  int rc;
  if ((rc = bad_policy(ctx, policy))) return rc;
  if (!ctx->staged) return fail(ctx, MI_ERR_NOT_STAGED, "mi_stage was not called");
  if (ctx->staged_total == 0) return MI_OK;
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  rc = run_kernel(ctx, ctx->staged_npos, ctx->staged_per_pos, ctx->staged_total, ctx->staged_offs,
                  seed, epoch, policy, false, stats);
  finish_stats(stats, t0);
  return rc;
}

int mi_fetch_winners(mi_ctx *ctx, int8_t *winners, uint64_t n) {
  if (!ctx || (!winners && n)) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (!ctx->staged) return fail(ctx, MI_ERR_NOT_STAGED, "mi_stage was not called");
  if (n > ctx->staged_total) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "n exceeds the staged batch");
  if (n == 0) return MI_OK;
  MI_HIP(ctx, hipMemcpyAsync(winners, ctx->winners.ptr, n, hipMemcpyDeviceToHost, ctx->stream));
  MI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MI_OK;
}

int mi_check_divmod(mi_ctx *ctx, const uint32_t *r, const uint32_t *n, uint32_t count,
                    uint32_t *q, uint32_t *rem) {
  if (!ctx) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  if (count == 0) return MI_OK;
  if (!r || !n || !q || !rem) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "null buffer");
  for (uint32_t i = 0; i < count; i++)
    if (n[i] == 0) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "zero divisor");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  DevBuf buf;
  const size_t bytes = sizeof(uint32_t) * (size_t)count;
  MI_HIP(ctx, hipMalloc(&buf.ptr, 4 * bytes));
  uint32_t *d = (uint32_t *)buf.ptr;
  int rc = MI_OK;
  hipError_t e = hipMemcpyAsync(d, r, bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d + count, n, bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = gpuai::launch_divmod(d, d + count, count, d + 2 * (size_t)count,
                                                d + 3 * (size_t)count, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(q, d + 2 * (size_t)count, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(rem, d + 3 * (size_t)count, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) rc = hip_fail(ctx, e, "divmod check");
  (void)hipFree(buf.ptr);
  return rc;
}

int mi_copy_calibration(mi_ctx *ctx, uint64_t bytes, uint32_t repeats, float *avg_ms) {
  if (!ctx || !avg_ms || bytes == 0 || bytes % 16 || repeats == 0) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  void *buf = nullptr;
  MI_HIP(ctx, hipMalloc(&buf, 2 * bytes));
  const uint4 *src = (const uint4 *)buf;
  uint4 *dst = (uint4 *)((char *)buf + bytes);
  hipError_t e = hipMemsetAsync(buf, 1, bytes, ctx->stream);
  float total = 0.f;
  for (uint32_t i = 0; i < repeats && e == hipSuccess; i++) {
    e = hipEventRecord(ctx->ev_start, ctx->stream);
    if (e == hipSuccess) e = gpuai::launch_copy(src, dst, bytes / 16, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev_stop, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev_stop);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    total += ms;
  }
  int rc = e == hipSuccess ? MI_OK : hip_fail(ctx, e, "copy calibration");
  (void)hipFree(buf);
  *avg_ms = total / (float)repeats;
  return rc;
}

int mi_valu_rate(mi_ctx *ctx, uint32_t iters, uint32_t repeats, double *lane_ops_per_s,
                 float *avg_ms) {
  if (!ctx || !lane_ops_per_s || !avg_ms || iters == 0 || iters > (1u << 20) || repeats == 0)
    return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  // 32 waves per CU (8 per SIMD), enough to issue a VALU instruction every cycle on each SIMD
  const uint32_t threads = ctx->cus * 32u * 64u;
  void *buf = nullptr;
  MI_HIP(ctx, hipMalloc(&buf, (size_t)threads * sizeof(uint32_t)));
  hipError_t e = hipSuccess;
  float total = 0.f, best = 0.f;
  for (uint32_t i = 0; i < repeats && e == hipSuccess; i++) {
    e = hipEventRecord(ctx->ev_start, ctx->stream);
    if (e == hipSuccess) e = gpuai::launch_valu_rate(iters, 0x5bd1e995u + i, (uint32_t *)buf, threads, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev_stop, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev_stop);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    total += ms;
    if (e == hipSuccess && (best == 0.f || ms < best)) best = ms;
  }
  int rc = e == hipSuccess ? MI_OK : hip_fail(ctx, e, "VALU rate probe");
  (void)hipFree(buf);
  *avg_ms = total / (float)repeats;
  *lane_ops_per_s = best > 0.f ? (double)threads * iters * 32.0 / (best * 1e-3) : 0.0;
  return rc;
}

// ---------------------------------------------------------------------------------------------
// graph-replayed batches
// ---------------------------------------------------------------------------------------------
struct mi_graph {
  mi_ctx *ctx = nullptr;
  uint32_t max_pos = 0, max_play = 0, policy = 0;
  // pinned host blob and its device twin: LaunchDesc | positions | offsets (npos + 1) | keys
  // (npos, keyed batches only), packed per batch so that one copy of the used prefix uploads it
  void *h_blob = nullptr, *d_blob = nullptr;
  size_t blob_bytes = 0;
  // pinned host output and its device twin: WorkWords | winners (max_play)
  void *h_out = nullptr, *d_out = nullptr;
  size_t out_bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphNode_t n_reset = nullptr, n_play = nullptr, n_d2h = nullptr;
  // the playout node's arguments and launch shape, fixed at creation except for the grid, which
  // each replay sizes to its batch (a small batch spread over every CU, as mi_run_playouts does)
  gpuai::PlayoutArgs args;
  uint32_t block = 0, resident = 0, grid = 0, refill = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // a replay launched by mi_graph_launch and not yet finished
  bool pending = false;
  uint64_t pending_total = 0;
};

namespace {

void graph_release(mi_graph *g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->e0) (void)hipEventDestroy(g->e0);
  if (g->e1) (void)hipEventDestroy(g->e1);
  if (g->d_blob) (void)hipFree(g->d_blob);
  if (g->d_out) (void)hipFree(g->d_out);
  if (g->h_blob) (void)hipHostFree(g->h_blob);
  if (g->h_out) (void)hipHostFree(g->h_out);
  delete g;
}

}  // namespace

int mi_graph_create(mi_ctx *ctx, uint32_t max_positions, uint32_t max_playouts, uint32_t policy,
                    mi_graph **out) {
  if (!ctx || !out) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  *out = nullptr;
  int rc;
  if ((rc = bad_policy(ctx, policy))) return rc;
  if (max_positions == 0 || max_playouts == 0 || max_playouts >= (1u << 31))
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "graph capacities must be in [1, 2^31)");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  mi_graph *g = new mi_graph();
  g->ctx = ctx;
  g->max_pos = max_positions;
  g->max_play = max_playouts;
  g->policy = policy;
  g->blob_bytes = sizeof(gpuai::LaunchDesc) + (sizeof(mi_pos) + 2 * sizeof(uint32_t)) *
                                                  (size_t)max_positions + sizeof(uint32_t);
  if (g->blob_bytes > 0xFFFFFFF0ull) {
    delete g;
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "too many positions for one graph batch");
  }
  g->out_bytes = sizeof(gpuai::WorkWords) + (size_t)max_playouts;
  hipError_t e = hipHostMalloc(&g->h_blob, g->blob_bytes, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&g->h_out, g->out_bytes, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&g->d_blob, g->blob_bytes);
  if (e == hipSuccess) e = hipMalloc(&g->d_out, g->out_bytes);
  if (e == hipSuccess) e = hipEventCreate(&g->e0);
  if (e == hipSuccess) e = hipEventCreate(&g->e1);
  if (e != hipSuccess) {
    rc = hip_fail(ctx, e, "graph buffers");
    graph_release(g);
    return rc;
  }
  std::memset(g->h_blob, 0, g->blob_bytes);
  gpuai::WorkWords *work = (gpuai::WorkWords *)g->d_out;
  gpuai::PlayoutArgs &a = g->args;
  std::memset(&a, 0, sizeof(a));
  a.per_pos_shift = 32u;  // graph batches always carry offsets
  // positions, offsets and keys are found through the description's section offsets
  a.pos = (const mi_pos *)((char *)g->d_blob + sizeof(gpuai::LaunchDesc));
  a.offs = nullptr;
  a.keys = nullptr;
  a.winners = (int8_t *)((char *)g->d_out + sizeof(gpuai::WorkWords));
  a.work = work;
  // the shape of a full batch; each replay takes its own batch's shape with this refill policy
  const mi_tuning t = shape_for(ctx, policy, max_playouts);
  a.defer_min = t.defer_min;
  a.desc = (const gpuai::LaunchDesc *)g->d_blob;
  g->block = t.block_size;
  g->refill = t.refill;
  g->resident = ctx->cus * t.blocks_per_cu;
  g->grid = grid_for(ctx, t, max_playouts);
  void *reset_args[] = {&work};
  void *play_args[] = {&a};
  hipKernelNodeParams rp;
  std::memset(&rp, 0, sizeof(rp));
  rp.func = const_cast<void *>(gpuai::reset_kernel());
  rp.gridDim = dim3(1);
  rp.blockDim = dim3(64);
  rp.kernelParams = reset_args;
  hipKernelNodeParams pp;
  std::memset(&pp, 0, sizeof(pp));
  pp.func = const_cast<void *>(gpuai::playout_kernel(policy, g->refill));
  pp.gridDim = dim3(g->grid);
  pp.blockDim = dim3(g->block);
  pp.kernelParams = play_args;
  e = hipGraphCreate(&g->graph, 0);
  // The batch's upload is not a node: an asynchronous copy on the stream before each replay
  // moves a 64k-leaf batch (1.3 MB) in ~40 us, a memcpy node took ~320 us.
  if (e == hipSuccess) e = hipGraphAddKernelNode(&g->n_reset, g->graph, nullptr, 0, &rp);
  if (e == hipSuccess) e = hipGraphAddKernelNode(&g->n_play, g->graph, &g->n_reset, 1, &pp);
  if (e == hipSuccess)
    e = hipGraphAddMemcpyNode1D(&g->n_d2h, g->graph, &g->n_play, 1, g->h_out, g->d_out,
                                g->out_bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    rc = hip_fail(ctx, e, "graph construction");
    graph_release(g);
    return rc;
  }
  *out = g;
  return MI_OK;
}

int mi_graph_launch(mi_graph *g, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                    const uint32_t *keys, uint32_t seed, uint32_t epoch) {
  if (!g) return MI_ERR_INVALID_ARGUMENT;
  mi_ctx *ctx = g->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  if (g->pending) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "the previous replay was not finished");
  if (npos > 0 && (!pos || !counts)) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "null batch");
  if (npos > g->max_pos) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "more positions than the graph holds");
  uint64_t total = 0;
  int rc;
  if ((rc = plan_batch(ctx, pos, npos, counts, 0, total))) return rc;
  if (total > g->max_play) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "more playouts than the graph holds");
  if (total > 0) {
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
    gpuai::LaunchDesc *d = (gpuai::LaunchDesc *)g->h_blob;
    const size_t pos_at = sizeof(gpuai::LaunchDesc);
    const size_t offs_at = pos_at + sizeof(mi_pos) * (size_t)npos;
    const size_t keys_at = offs_at + sizeof(uint32_t) * ((size_t)npos + 1);
    const size_t in_bytes = keys_at + (keys ? sizeof(uint32_t) * (size_t)npos : 0);
    std::memset(d, 0, sizeof(*d));
    d->npos = npos;
    d->total = (uint32_t)total;
    d->seed = seed;
    d->epoch = epoch;
    d->ordered = 0;
    d->pos_at = (uint32_t)pos_at;
    d->offs_at = (uint32_t)offs_at;
    d->keys_at = keys ? (uint32_t)keys_at : 0u;
    char *h = (char *)g->h_blob;
    std::memcpy(h + pos_at, pos, sizeof(mi_pos) * (size_t)npos);
    std::memcpy(h + offs_at, ctx->host_offs.data(), sizeof(uint32_t) * ((size_t)npos + 1));
    if (keys) std::memcpy(h + keys_at, keys, sizeof(uint32_t) * (size_t)npos);
    const size_t out_bytes = sizeof(gpuai::WorkWords) + (size_t)total;
    MI_HIP(ctx, hipGraphExecMemcpyNodeSetParams1D(g->exec, g->n_d2h, g->h_out, g->d_out, out_bytes,
                                                  hipMemcpyDeviceToHost));
    // the batch's shape from the tuning table (the graph's refill policy; both are correct for
    // any grid: static refill strides by the grid's threads)
    const mi_tuning t = shape_for(ctx, g->policy, total, (int)g->refill);
    const uint64_t need = (total + t.block_size - 1u) / t.block_size;
    const uint32_t grid = gpuai::grid_size(need, (uint64_t)ctx->cus * t.blocks_per_cu);
    if (grid != g->grid || t.block_size != g->block || t.defer_min != g->args.defer_min) {
      g->args.defer_min = t.defer_min;
      void *play_args[] = {&g->args};
      hipKernelNodeParams pp;
      std::memset(&pp, 0, sizeof(pp));
      pp.func = const_cast<void *>(gpuai::playout_kernel(g->policy, g->refill));
      pp.gridDim = dim3(grid);
      pp.blockDim = dim3(t.block_size);
      pp.kernelParams = play_args;
      MI_HIP(ctx, hipGraphExecKernelNodeSetParams(g->exec, g->n_play, &pp));
      g->grid = grid;
      g->block = t.block_size;
      g->resident = ctx->cus * t.blocks_per_cu;
    }
    MI_HIP(ctx, hipEventRecord(g->e0, ctx->stream));
    MI_HIP(ctx, hipMemcpyAsync(g->d_blob, g->h_blob, in_bytes, hipMemcpyHostToDevice, ctx->stream));
    MI_HIP(ctx, hipGraphLaunch(g->exec, ctx->stream));
    MI_HIP(ctx, hipEventRecord(g->e1, ctx->stream));
  }
  g->pending = true;
  g->pending_total = total;
  return MI_OK;
}

int mi_graph_finish(mi_graph *g, int8_t *winners, mi_stats *stats) {
  if (!g) return MI_ERR_INVALID_ARGUMENT;
  mi_ctx *ctx = g->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (!g->pending) return fail(ctx, MI_ERR_NOT_STAGED, "no replay was launched");
  const uint64_t total = g->pending_total;
  if (total > 0 && !winners) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "winners pointer is null");
  g->pending = false;  // the replay is consumed whatever happens below
  if (total == 0) return MI_OK;
  MI_HIP(ctx, hipEventSynchronize(g->e1));
  std::memcpy(winners, (const char *)g->h_out + sizeof(gpuai::WorkWords), (size_t)total);
  if (stats) {
    const gpuai::WorkWords *w = (const gpuai::WorkWords *)g->h_out;
    float ms = 0.f;
    MI_HIP(ctx, hipEventElapsedTime(&ms, g->e0, g->e1));
    stats->playouts = total;
    stats->plies = w->plies;
    stats->kernel_ms = ms;
    stats->max_moves = w->max_moves;
    stats->grid_blocks = g->grid;
    stats->block_size = g->block;
    stats->refill = g->refill;
  }
  return MI_OK;
}

int mi_graph_run(mi_graph *g, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                 const uint32_t *keys, uint32_t seed, uint32_t epoch, int8_t *winners,
                 mi_stats *stats) {
  const auto t0 = Clock::now();
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (npos > 0 && counts && pos) {  // a null winners buffer is refused before anything runs
    uint64_t total = 0;
    for (uint32_t i = 0; i < npos; i++) total += counts[i];
    if (total > 0 && !winners) return g ? fail(g->ctx, MI_ERR_INVALID_ARGUMENT, "winners pointer is null")
                                        : MI_ERR_INVALID_ARGUMENT;
  }
  int rc = mi_graph_launch(g, pos, npos, counts, keys, seed, epoch);
  if (rc) return rc;
  rc = mi_graph_finish(g, winners, stats);
  finish_stats(stats, t0);
  return rc;
}

int mi_graph_destroy(mi_graph *g) {
  if (!g) return MI_OK;
  std::lock_guard<std::mutex> lock(g->ctx->mu);
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  graph_release(g);
  return MI_OK;
}

}  // extern "C"
