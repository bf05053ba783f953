// kernels.h -- host-side launch interface of the gfx950 kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpuai/mi_playout.h"

namespace gpuai {

// The queue's playout indices are dealt to slabs in blocks of kSlabBlock: with S = min(kSlabs,
// grid) slabs, block b (indices 128b .. 128b+127) belongs to slab b % S, and the waves of
// workgroup w claim from slab w % S only.  Workgroups are dispatched round-robin over the chip's
// eight XCDs, so a slab's blocks -- and the 128-byte lines of winners they fill -- stay with one
// XCD's L2 (16 slabs: two per XCD); each slab's counter sits alone in a 128-byte line, so claims
// of different slabs do not contend for one line (a single shared counter throttled the claims:
// 4-6% of a launch).  Grids
// are a multiple of S (grid_size), so every slab is served by as many workgroups as any other
// and holds the same share of the work, give or take one block; a wave whose slab is dry is
// done.
constexpr uint32_t kSlabs = 16;
constexpr uint32_t kSlabBlock = 128;

// Device-side per-launch work words (zeroed by launch_reset before use).
struct WorkWords {
  unsigned long long plies;  // plies applied over the launch
  uint32_t max_moves;        // longest move list seen
  uint32_t overflow;         // a move list exceeded a capacity
  uint32_t pad0[28];
  struct alignas(128) QueueLine {
    uint32_t next;  // next unclaimed slab-local index (wave-level claims)
    uint32_t pad[31];
  } queue[kSlabs];
};
static_assert(sizeof(WorkWords) == 128 * (1 + kSlabs), "header line + one line per slab");

// Workgroups of a persistent launch for `need` workgroups' worth of work when `resident` fit on
// the chip: a multiple of the slab count above it.
inline uint32_t grid_size(uint64_t need, uint64_t resident) {
  uint64_t g = need < resident ? need : resident;
  if (g > kSlabs) g = (g + kSlabs - 1) / kSlabs * kSlabs;
  return (uint32_t)(g ? g : 1);
}

// Batch description read by the kernel from device memory (graph replays change it between
// launches without touching the captured kernel arguments).
struct LaunchDesc {
  uint32_t npos;
  uint32_t total;
  uint32_t seed;
  uint32_t epoch;
  uint32_t ordered;  // 0: the hand-out order of the arguments is ignored (positions in index order)
  // Byte offsets from this description of the batch's sections when they follow it in one
  // block of device memory (graph replays pack positions, offsets and keys behind it, so that one
  // copy of the used prefix uploads a batch); 0: the pointer of the arguments (keys: none).
  uint32_t pos_at, offs_at, keys_at;
  uint32_t pad[8];
};

struct PlayoutArgs {
  const mi_pos *pos;       // npos positions
  const uint32_t *offs;    // npos+1 prefix offsets of playout counts, or null (uniform per_pos)
  // Optional hand-out order.  When `order` is set, queue index g is taken in rank space:
  // rank r = the position handed out r-th, order[r] its index in `pos`, ord_pos[r] its copy,
  // ord_offs (uneven counts) the prefix offsets in rank order.  Results still go to the
  // position's own slots, so the order changes nothing but the schedule.
  const mi_pos *ord_pos;
  const uint32_t *order;
  const uint32_t *ord_offs;
  // Optional stream keys: the Philox key of position i's playouts is (seed, keys[i]) instead of
  // (seed, i), so that a playout's stream need not depend on where its position sits in a batch.
  const uint32_t *keys;
  uint32_t npos;
  uint32_t per_pos;
  uint32_t per_pos_shift;  // log2(per_pos) when per_pos is a power of two, else 32
  uint32_t total;          // total playouts (< 2^31)
  uint32_t seed;
  uint32_t epoch;
  int8_t *winners;         // total
  uint32_t *plies;         // total (trace mode only)
  uint32_t *hashes;        // total (trace mode only)
  WorkWords *work;
  uint32_t defer_min;      // deferred multi-jump lanes that trigger a walk phase
  const LaunchDesc *desc;  // when set, npos / total / seed / epoch come from here
};

// Kernel entry of the playout kernel for a policy and refill policy (graph construction).
const void *playout_kernel(uint32_t policy, uint32_t refill);
const void *reset_kernel();

hipError_t launch_reset(WorkWords *work, hipStream_t stream);
hipError_t launch_playouts(const PlayoutArgs &args, uint32_t policy, bool trace, uint32_t refill,
                           uint32_t grid, uint32_t block, hipStream_t stream);
hipError_t launch_gen_moves(const mi_pos *pos, uint32_t npos, mi_move *moves, uint32_t cap,
                            uint32_t *counts, WorkWords *work, hipStream_t stream);
hipError_t launch_start_positions(uint32_t seed, uint32_t n, uint32_t kmin, uint32_t kmax,
                                  uint32_t reject_terminal, mi_pos *out, WorkWords *work,
                                  hipStream_t stream);
hipError_t launch_divmod(const uint32_t *r, const uint32_t *n, uint32_t count, uint32_t *q,
                         uint32_t *rem, hipStream_t stream);
hipError_t launch_copy(const uint4 *src, uint4 *dst, uint64_t n16, hipStream_t stream);
// `threads` must be a multiple of 256; `out` holds `threads` words.  32 VALU instructions per
// lane per iteration.
hipError_t launch_valu_rate(uint32_t iters, uint32_t salt, uint32_t *out, uint32_t threads,
                            hipStream_t stream);
hipError_t launch_tally(const int8_t *winners, const uint32_t *offs, uint32_t npos,
                        uint32_t per_pos, uint32_t *tallies, hipStream_t stream);

}  // namespace gpuai
