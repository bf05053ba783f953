// noalloc.hip -- the allocation-free entry point mi_run_playouts_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture, serving loops that
// forbid hipMalloc) hand in positions, counts and the winners buffer in device memory; what the
// playout kernel needs beyond them -- its work words, the batch description, playout offsets and
// the longest-first hand-out order -- lives in `scratch`, a file-scope device array of fixed size
// (each device holds its own copy until the process ends).  Host code never names the array in a
// kernel argument, copy or fill: noalloc_setup() looks up its device address once per device
// (hipGetSymbolAddress, with the device current) and every launch takes that address.
//
// One call is two kernels on the caller's stream: k_noalloc_prep (one workgroup: validation,
// work-word reset, offsets, counting sort by piece count) and the persistent playout kernel,
// which reads the batch from the scratch's LaunchDesc.  Nothing waits on the host.  Every call
// on a device shares that device's scratch, and each call rewrites all of it that the next
// kernel reads, so consecutive calls on one stream are independent; calls that overlap in time
// on one device (different streams or threads) race on the scratch.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ctx.h"
#include "gpuai/mi_playout.h"
#include "kernels.h"

namespace gpuai {

namespace {

constexpr uint32_t kMaxPos = MI_NOALLOC_MAX_POSITIONS;
constexpr uint32_t kPrepThreads = 256;
constexpr uint32_t kBuckets = 33;  // piece counts 0..32

// Layout of the scratch array.
struct Scratch {
  WorkWords work;
  LaunchDesc desc;
  uint32_t offs[kMaxPos + 1];      // prefix offsets of the counts, index order
  uint32_t ord_offs[kMaxPos + 1];  // prefix offsets in hand-out order
  uint32_t order[kMaxPos];         // hand-out rank -> position index
  mi_pos ord_pos[kMaxPos];         // positions in hand-out order
};
static_assert(sizeof(LaunchDesc) == 64, "64-byte batch description");

static __device__ __align__(256) unsigned char scratch[sizeof(Scratch)];

constexpr int kMaxDevices = 64;
std::mutex g_setup_mu;
void *g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// Exclusive prefix sum over the workgroup of one 64-bit value per thread (kPrepThreads threads,
// four waves); `total` receives the sum.  Every thread of the workgroup must call it.
__device__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *lds_waves, uint64_t &total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint64_t incl = v;
  for (uint32_t off = 1; off < 64u; off <<= 1) {
    const uint64_t o = __shfl_up(incl, off);
    if (lane >= off) incl += o;
  }
  if (lane == 63u) lds_waves[wave] = incl;
  __syncthreads();
  uint64_t before = 0, sum = 0;
  for (uint32_t w = 0; w < kPrepThreads / 64u; w++) {
    const uint64_t x = lds_waves[w];
    before += w < wave ? x : 0ull;
    sum += x;
  }
  __syncthreads();  // lds_waves may be reused by the next call
  total = sum;
  return before + incl - v;
}

// One workgroup of kPrepThreads threads; thread t owns positions [t*chunk, (t+1)*chunk).
// status (may be null) receives MI_OK, or MI_ERR_INVALID_ARGUMENT when a position is malformed
// (overlapping sides, turn > 1) or the counts add up to 2^31 or more; the batch then plays
// nothing.
__global__ void __launch_bounds__(kPrepThreads)
    k_noalloc_prep(const mi_pos *__restrict__ pos, uint32_t npos,
                   const uint32_t *__restrict__ counts, uint32_t per_pos, uint32_t seed,
                   uint32_t epoch, uint32_t order_min_total, Scratch *S, uint32_t *status) {
  __shared__ uint32_t hist[kBuckets][kPrepThreads];
  __shared__ uint64_t waves[kPrepThreads / 64u];
  __shared__ uint32_t any_bad;
  const uint32_t t = threadIdx.x;
  const uint32_t chunk = (npos + kPrepThreads - 1u) / kPrepThreads;
  const uint32_t lo = min(t * chunk, npos), hi = min(lo + chunk, npos);
  if (t == 0u) any_bad = 0u;
  for (uint32_t b = 0; b < kBuckets; b++) hist[b][t] = 0u;
  __syncthreads();
  // validation, piece-count histogram and this thread's playout total
  uint32_t bad = 0u;
  uint64_t mine = 0;
  for (uint32_t i = lo; i < hi; i++) {
    const mi_pos q = pos[i];
    bad |= ((q.p1 & q.p2) != 0u) | (q.turn > 1u);
    hist[32u - (uint32_t)__popc(q.p1 | q.p2)][t]++;
    mine += counts ? (uint64_t)counts[i] : (uint64_t)per_pos;
  }
  if (bad) atomicOr(&any_bad, 1u);
  uint64_t total = 0;
  const uint64_t before = block_exclusive_scan(mine, waves, total);
  __syncthreads();
  const bool ok = any_bad == 0u && total < (1ull << 31);
  if (ok && counts) {  // offsets in index order
    uint32_t run = (uint32_t)before;
    for (uint32_t i = lo; i < hi; i++) {
      S->offs[i] = run;
      run += counts[i];
    }
    if (t == kPrepThreads - 1u) S->offs[npos] = (uint32_t)total;
  }
  // Longest-first hand-out (positions with more pieces first; stable), as mi_run_playouts
  // orders batches of at least 256 playouts per compute unit.  The histogram, read bucket-major
  // (bucket b, then thread), becomes each thread's first rank in each bucket.
  const bool ordered = ok && npos > 1u && total >= order_min_total;
  if (ordered) {
    uint32_t *flat = &hist[0][0];
    const uint32_t per = kBuckets;  // kBuckets * kPrepThreads entries, kBuckets per thread
    uint64_t s = 0;
    for (uint32_t k = 0; k < per; k++) s += flat[t * per + k];
    uint64_t unused = 0;
    uint32_t run = (uint32_t)block_exclusive_scan(s, waves, unused);
    for (uint32_t k = 0; k < per; k++) {
      const uint32_t c = flat[t * per + k];
      flat[t * per + k] = run;
      run += c;
    }
    __syncthreads();
    for (uint32_t i = lo; i < hi; i++) {
      const mi_pos q = pos[i];
      const uint32_t r = hist[32u - (uint32_t)__popc(q.p1 | q.p2)][t]++;
      S->order[r] = i;
      S->ord_pos[r] = q;
      if (counts) S->ord_offs[r + 1u] = counts[i];  // the counts in rank order, summed below
    }
    __syncthreads();  // workgroup-scope fence: the scattered counts are read by other threads
    if (counts) {
      uint64_t c = 0;
      for (uint32_t r = lo; r < hi; r++) c += S->ord_offs[r + 1u];
      uint64_t unused2 = 0;
      uint32_t acc = (uint32_t)block_exclusive_scan(c, waves, unused2);
      for (uint32_t r = lo; r < hi; r++) {
        acc += S->ord_offs[r + 1u];
        S->ord_offs[r + 1u] = acc;
      }
      if (t == 0u) S->ord_offs[0] = 0u;
    }
  }
  if (t == 0u) {
    for (uint32_t q = 0; q < kSlabs; q++) S->work.queue[q].next = 0u;
    S->work.max_moves = 0u;
    S->work.overflow = 0u;
    S->work.plies = 0ull;
    S->desc.npos = npos;
    S->desc.total = ok ? (uint32_t)total : 0u;
    S->desc.seed = seed;
    S->desc.epoch = epoch;
    S->desc.ordered = ordered ? 1u : 0u;
    S->desc.pos_at = 0u;  // positions, offsets and order are the kernel arguments
    S->desc.offs_at = 0u;
    S->desc.keys_at = 0u;
    if (status) *status = ok ? (uint32_t)MI_OK : (uint32_t)MI_ERR_INVALID_ARGUMENT;
  }
}

}  // namespace

// Looks up this device's address of `scratch` (once per device; the device must be current).
int noalloc_setup(mi_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return ctxutil::fail(ctx, MI_ERR_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_setup_mu);
  if (g_scratch_ptr[ctx->device]) return MI_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return ctxutil::hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[ctx->device] = ptr;
  return MI_OK;
}

}  // namespace gpuai

extern "C" int mi_run_playouts_noalloc(mi_ctx *ctx, const mi_pos *pos, uint32_t npos,
                                       const uint32_t *counts, uint32_t per_pos, uint32_t seed,
                                       uint32_t epoch, uint32_t policy, int8_t *winners,
                                       uint32_t *status, void *stream) {
  using gpuai::ctxutil::fail;
  if (!ctx) return MI_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  if (policy > MI_POLICY_HEURISTIC_NOISY) return fail(ctx, MI_ERR_INVALID_ARGUMENT, "unknown policy");
  if (npos > gpuai::kMaxPos)
    return fail(ctx, MI_ERR_INVALID_ARGUMENT,
                "mi_run_playouts_noalloc takes at most 65536 positions (its fixed scratch size)");
  if (npos > 0 && (!pos || !winners))
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "positions or winners pointer is null");
  if (!counts && (uint64_t)npos * per_pos >= (1ull << 31))
    return fail(ctx, MI_ERR_INVALID_ARGUMENT, "more than 2^31-1 playouts in one call");
  void *ptr = (ctx->device >= 0 && ctx->device < gpuai::kMaxDevices)
                  ? gpuai::g_scratch_ptr[ctx->device]
                  : nullptr;
  if (!ptr) return fail(ctx, MI_ERR_DEVICE, "the scratch of this device was not set up");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, MI_ERR_DEVICE, "hipSetDevice failed");
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  gpuai::Scratch *S = (gpuai::Scratch *)ptr;
  const uint32_t order_min_total = ctx->cus * 256u;
  hipLaunchKernelGGL(gpuai::k_noalloc_prep, dim3(1), dim3(gpuai::kPrepThreads), 0, s, pos, npos,
                     counts, per_pos, seed, epoch, order_min_total, S, status);
  MI_HIP(ctx, hipGetLastError());
  const uint64_t uniform_total = counts ? 0ull : (uint64_t)npos * per_pos;
  if (npos == 0 || (!counts && uniform_total == 0)) return MI_OK;
  gpuai::PlayoutArgs a;
  a.pos = pos;
  a.offs = counts ? S->offs : nullptr;
  a.ord_pos = S->ord_pos;  // used only when the prep kernel ordered the batch (desc.ordered)
  a.order = S->order;
  a.ord_offs = counts ? S->ord_offs : nullptr;
  a.keys = nullptr;
  a.npos = npos;
  a.per_pos = per_pos;
  a.per_pos_shift = per_pos && !(per_pos & (per_pos - 1u)) ? (uint32_t)__builtin_ctz(per_pos) : 32u;
  a.total = 0;  // read from the description
  a.seed = seed;
  a.epoch = epoch;
  a.winners = winners;
  a.plies = nullptr;
  a.hashes = nullptr;
  a.work = &S->work;
  a.defer_min = ctx->tuning.defer_min;
  a.desc = &S->desc;
  // the grid as mi_run_playouts sizes it; with counts in device memory the total is unknown
  // here, so the whole resident grid runs (waves without work leave at their first refill)
  const uint64_t block = ctx->tuning.block_size;
  const uint64_t resident = (uint64_t)ctx->cus * ctx->tuning.blocks_per_cu;
  const uint64_t need = counts ? resident : (uniform_total + block - 1) / block;
  const uint32_t grid = gpuai::grid_size(need, resident);
  MI_HIP(ctx, gpuai::launch_playouts(a, policy, false, ctx->tuning.refill, grid, (uint32_t)block,
                                     s));
  return MI_OK;
}
