// playout_kernels.hip -- gfx950 kernels of the playout engine.
//
// k_playouts is a persistent kernel: each lane of a wave64 runs one playout at a time with the
// position in VGPRs (ck::Lane) and, when its game ends, takes the next playout index.  New
// playouts start only on wave iterations that are multiples of four, so every active lane is at
// a ply index that is a multiple of four on those iterations and the Philox4x32-10 call that
// yields the next four words is wave-uniform (one call per four plies instead of a divergent
// call on almost every iteration).  Refill policy MI_REFILL_QUEUE claims playout indices for the
// whole wave in multiples of kChunk from its slab's device counter (one atomic per claim,
// broadcast with a lane read; slabs are described in kernels.h), MI_REFILL_STATIC strides lanes
// over the index space without atomics.
#include <hip/hip_runtime.h>

#include "checkers.h"
#include "kernels.h"

namespace gpuai {

namespace {

// Queue claims are made in multiples of kChunk playouts (a wave holds at most one partly used
// claim).  With one counter per slab in a line of its own, claims of 16 beat 32 by about 1%
// (less unplayed work held when a slab runs dry); on a single shared counter, 16 had lost 9% to
// the contention of twice as many atomics on one line.  Each claim is a device atomic, which
// travels to memory as a 32-byte request (random-1024x1024: about 65,500 per launch, twice the
// winners' bytes in the write-request counters).  Workgroup pools in LDS, refilled by guided
// claims from the slab counter, cut that to about 5,000 but cost 5-9% of the launch in
// instructions on the refill path, which this VALU-bound loop cannot hide; claims stay per wave.
constexpr uint32_t kChunk = 16;

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__global__ void k_reset(WorkWords *work) {
  if (threadIdx.x < kSlabs) work->queue[threadIdx.x].next = 0;
  if (threadIdx.x == 0) {
    work->max_moves = 0;
    work->overflow = 0;
    work->plies = 0ull;
  }
}

// Queue index of slab-local index c of slab s of ns (blocks of kSlabBlock dealt round-robin).
__device__ __forceinline__ uint32_t slab_index(uint32_t c, uint32_t s, uint32_t ns) {
  return (((c / kSlabBlock) * ns + s) * kSlabBlock) | (c % kSlabBlock);
}

// Position p, playout j and result slot `out` of the playout with queue index g, and the
// position itself.
__device__ __forceinline__ void locate(const PlayoutArgs &a, uint32_t g, uint32_t &p, uint32_t &j,
                                       uint32_t &out, mi_pos &q) {
  const uint32_t *offs = a.order ? a.ord_offs : a.offs;
  uint32_t r;
  if (offs) {
    uint32_t lo = 0, hi = a.npos;
    while (hi - lo > 1u) {
      uint32_t mid = (lo + hi) >> 1;
      if (offs[mid] <= g) lo = mid; else hi = mid;
    }
    r = lo;
    j = g - offs[lo];
  } else if (a.per_pos_shift < 32u) {
    r = g >> a.per_pos_shift;
    j = g & (a.per_pos - 1u);
  } else {
    ck::divmod_u32(g, a.per_pos, r, j);
  }
  if (a.order) {  // two independent loads
    p = a.order[r];
    q = a.ord_pos[r];
    out = a.offs ? a.offs[p] + j : p * a.per_pos + j;
  } else {
    p = r;
    q = a.pos[r];
    out = g;
  }
}

// Philox words of a lane: `cur` holds block `blk` of its stream (words 4*blk .. 4*blk+3),
// `nxt` block blk+1 when nxt_ok.  Blocks are generated only on wave iterations that are
// multiples of four (wave-uniform work); since a lane consumes at most one word per iteration,
// holding the current and the next block always covers the next four iterations.
struct Words {
  uint32_t c0, c1, c2, c3, n0, n1, n2, n3;
  uint32_t blk;
  bool cur_ok, nxt_ok;
};

__device__ __forceinline__ uint32_t take_word(Words &w, uint32_t ply) {
  if ((ply >> 2) != w.blk) {
    w.c0 = w.n0;
    w.c1 = w.n1;
    w.c2 = w.n2;
    w.c3 = w.n3;
    w.blk++;
    w.nxt_ok = false;
  }
  const uint32_t q = ply & 3u;
  return q == 0u ? w.c0 : q == 1u ? w.c1 : q == 2u ? w.c2 : w.c3;
}

template <uint32_t POLICY, bool TRACE, uint32_t REFILL>
__global__ void __launch_bounds__(1024) k_playouts(PlayoutArgs a) {
  if (a.desc) {  // graph replays and allocation-free launches: the batch lives in device memory
    const LaunchDesc *d = a.desc;
    a.npos = d->npos;
    a.total = d->total;
    a.seed = d->seed;
    a.epoch = d->epoch;
    if (!d->ordered) a.order = nullptr;
    const char *base = (const char *)d;
    if (d->pos_at) a.pos = (const mi_pos *)(base + d->pos_at);
    if (d->offs_at) a.offs = (const uint32_t *)(base + d->offs_at);
    a.keys = d->keys_at ? (const uint32_t *)(base + d->keys_at) : nullptr;
  }
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t nthreads = gridDim.x * blockDim.x;
  bool active = false, deferred = false;
  uint32_t g = 0, p = 0, j = 0, ply = 0, h = 0, maxn = 0;
  Words w = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, false, false};
  uint64_t plies_acc = 0;
  ck::Lane L = {0u, 0u, 0u, 0u, 0u};
  ck::Gen gd;  // the analysis of a deferred ply, kept for its walk (saves a second analyse)
  // wave-uniform refill and deferral state: the reservation [res_next, res_end) is slab-local in
  // slab `slab`
  uint32_t res_next = 0, res_end = 0, since_walk = 0;
  const uint32_t nslabs = gridDim.x < kSlabs ? gridDim.x : kSlabs;
  const uint32_t slab = blockIdx.x % nslabs;
  bool exhausted = false;
  uint32_t static_next = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t defer_min = a.defer_min;

  for (uint32_t iter = 0;; iter++) {
    if ((iter & 3u) == 0u) {
      bool start = false;
      if (REFILL == MI_REFILL_STATIC) {
        if (!active && static_next < a.total) {
          g = static_next;
          static_next += nthreads;
          start = true;
        }
      } else if (!exhausted) {
        const uint64_t idle = __ballot(!active);
        if (idle) {
          const uint32_t cnt = (uint32_t)__popcll(idle);
          const uint32_t rank = lane_rank(idle);  // of an idle lane among the idle lanes
          const uint32_t avail = res_end - res_next;
          uint32_t c = 0;
          // the new claim covers every idle lane that the old one cannot
          const uint32_t want = avail < cnt ? ((cnt - avail + kChunk - 1u) / kChunk) * kChunk : 0u;
          if (avail < cnt) {
            if (lane == 0u) c = atomicAdd(&a.work->queue[slab].next, want);
            c = __shfl(c, 0);
          }
          if (!active) {
            const uint32_t idx =
                slab_index(rank < avail ? res_next + rank : c + (rank - avail), slab, nslabs);
            if (idx < a.total) {
              g = idx;
              start = true;
            }
          }
          if (avail < cnt) {
            res_next = c + (cnt - avail);
            res_end = c + want;
          } else {
            res_next += cnt;
          }
          if (slab_index(res_next, slab, nslabs) >= a.total) exhausted = true;  // slab dry
        }
      }
      if (start) {
        active = true;
        deferred = false;
        mi_pos q;
        locate(a, g, p, j, g, q);  // from here on g is the playout's result slot
        if (a.keys) p = a.keys[p];  // from here on p is the stream key
        L = ck::lane_from_abs(q.p1, q.p2, q.kings, q.turn, q.counter);
        ply = 0;
        h = 0;
        w.cur_ok = false;
        w.nxt_ok = false;
      }
      if (!__any(active)) break;
      if (active && (!w.cur_ok || !w.nxt_ok)) {
        const uint32_t b = w.cur_ok ? w.blk + 1u : (ply >> 2);
        uint32_t o[4];
        ck::philox(b, j, ck::TAG_PLAYOUT, a.epoch, a.seed, p, o);
        if (w.cur_ok) {
          w.n0 = o[0];
          w.n1 = o[1];
          w.n2 = o[2];
          w.n3 = o[3];
          w.nxt_ok = true;
        } else {
          w.c0 = o[0];
          w.c1 = o[1];
          w.c2 = o[2];
          w.c3 = o[3];
          w.blk = b;
          w.cur_ok = true;
        }
      }
    }
    // common part of the ply; multi-jump plies of the random policy wait for the walk phase
    bool play = false;
    ck::PlyResult res;
    uint32_t ld = 0;
    if (active && !deferred) {
      const uint32_t r = take_word(w, ply);
      // a deferred ply takes the same word again in the walk phase (take_word is idempotent
      // for a given ply)
      bool d = false;
      if (POLICY == MI_POLICY_RANDOM) res = ck::ply_random_fast<TRACE>(L, r, &ld, d, gd);
      else if (POLICY == MI_POLICY_HEURISTIC) res = ck::ply_heuristic_fast<TRACE>(L, r, &ld, d, gd);
      else res = ck::ply_noisy_fast<TRACE>(L, r, &ld, d, gd);
      deferred = d;
      play = !d;
    }
    {
      const uint64_t dm = __ballot(deferred);
      if (dm) {
        const uint32_t nd = (uint32_t)__popcll(dm);
        const uint32_t na = (uint32_t)__popcll(__ballot(active));
        if (nd >= defer_min || nd == na || since_walk >= 8u) {
          since_walk = 0;
          if (deferred) {
            const uint32_t r = take_word(w, ply);
            if (POLICY == MI_POLICY_RANDOM) {
              // capture trees in closed form; the walk only for king trees past two jumps
              bool deep = false;
              res = ck::ply_random_trees<TRACE>(L, gd, r, &ld, deep);
              if (deep) res = ck::ply_random_given<TRACE>(L, gd, r, &ld);
            }
            else if (POLICY == MI_POLICY_HEURISTIC) res = ck::ply_heuristic_given<TRACE>(L, gd, r, &ld);
            else res = ck::ply_noisy_given<TRACE>(L, gd, r, &ld);
            deferred = false;
            play = true;
          }
        } else {
          since_walk++;
        }
      }
    }
    if (play) {
      if (TRACE) h = ck::hash_step(h, ck::ply_record(ld, res.n, res.chosen));
      maxn = res.n > maxn ? res.n : maxn;
      if (res.over) {
        a.winners[g] = (int8_t)res.winner;
        if (TRACE) {
          if (a.plies) a.plies[g] = ply;
          if (a.hashes) a.hashes[g] = h;
        }
        plies_acc += ply;
        active = false;
      } else {
        ply++;
      }
    }
  }
  // statistics: per wave, then per workgroup in LDS, then one device atomic each (device atomics
  // travel to memory as 32-byte requests)
  for (int off = 32; off >= 1; off >>= 1) {
    plies_acc += __shfl_xor(plies_acc, off);
    uint32_t m = __shfl_xor(maxn, off);
    maxn = m > maxn ? m : maxn;
  }
  __shared__ unsigned long long wg_plies;
  __shared__ uint32_t wg_maxn;
  if (threadIdx.x == 0) {
    wg_plies = 0ull;
    wg_maxn = 0u;
  }
  __syncthreads();
  if (lane == 0u) {
    atomicAdd(&wg_plies, (unsigned long long)plies_acc);
    atomicMax(&wg_maxn, maxn);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&a.work->plies, wg_plies);
    atomicMax(&a.work->max_moves, wg_maxn);
  }
}

// One lane per position: the whole move list in reference order, absolute squares.
__global__ void __launch_bounds__(256) k_gen_moves(const mi_pos *__restrict__ pos, uint32_t npos,
                                                   mi_move *__restrict__ out, uint32_t cap,
                                                   uint32_t *__restrict__ counts,
                                                   WorkWords *work) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npos) return;
  const mi_pos q = pos[i];
  const ck::Lane L = ck::lane_from_abs(q.p1, q.p2, q.kings, q.turn, q.counter);
  ck::Gen gen;
  ck::analyse(L, gen);
  const uint32_t n = gen.capture ? ck::count_walks(L, gen) : ck::slot_count(gen.slot);
  counts[i] = n;
  if (n > cap) {
    atomicOr(&work->overflow, 1u);
    return;
  }
  mi_move *dst = out + (size_t)i * cap;
  auto v = [&](uint32_t idx, uint32_t from, uint32_t to, uint32_t, uint32_t jumps, uint32_t dirs,
               uint32_t king) {
    const uint32_t ref = L.turn ? n - 1u - idx : idx;
    mi_move m;
    m.from = (uint8_t)ck::abs_sq(L, from);
    m.to = (uint8_t)ck::abs_sq(L, to);
    m.jumps = (uint8_t)jumps;
    m.promoted = (uint8_t)((!king && to >= ck::FAR_ROW_FIRST) ? 1 : 0);
    uint32_t x = from;
#pragma unroll
    for (uint32_t t = 0; t < 8u; t++) {
      m.inter[t] = 0;
      m.removed[t] = 0;
    }
    for (uint32_t t = 0; t < jumps; t++) {
      const uint32_t d = (dirs >> (2u * t)) & 3u;
      const uint32_t land = (uint32_t)((int)x + ck::jump_off(d));
      m.inter[t] = (uint8_t)ck::abs_sq(L, land);
      m.removed[t] = (uint8_t)ck::abs_sq(L, ck::mid_sq(x, land));
      x = land;
    }
    dst[ref] = m;
  };
  ck::enumerate(L, gen, v);
}

// One wave per position: tallies of draws, player-1 wins and player-2 wins.
__global__ void __launch_bounds__(256) k_tally(const int8_t *__restrict__ winners,
                                               const uint32_t *__restrict__ offs, uint32_t npos,
                                               uint32_t per_pos, uint32_t *__restrict__ tallies) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63u;
  if (wave >= npos) return;
  const uint64_t begin = offs ? offs[wave] : (uint64_t)wave * per_pos;
  const uint64_t end = offs ? offs[wave + 1] : begin + per_pos;
  uint32_t c[3] = {0u, 0u, 0u};
  for (uint64_t g = begin + lane; g < end; g += 64u) {
    const int w = winners[g];
    c[0] += w < 0;
    c[1] += w == 0;
    c[2] += w == 1;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    c[0] += __shfl_xor(c[0], off);
    c[1] += __shfl_xor(c[1], off);
    c[2] += __shfl_xor(c[2], off);
  }
  if (lane == 0u) {
    tallies[3u * wave + 0u] = c[0];
    tallies[3u * wave + 1u] = c[1];
    tallies[3u * wave + 2u] = c[2];
  }
}

// One lane per start position (the stream layout of mi_gen_start_positions).
__global__ void __launch_bounds__(256) k_start_positions(uint32_t seed, uint32_t n, uint32_t kmin,
                                                         uint32_t kmax, uint32_t reject,
                                                         mi_pos *__restrict__ out,
                                                         WorkWords *work) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ck::Lane L = ck::lane_from_abs(0x00000FFFu, 0xFFF00000u, 0u, 0u, 0u);
  bool done = false;
  for (uint32_t attempt = 0; attempt < 4096u && !done; attempt++) {
    L = ck::lane_from_abs(0x00000FFFu, 0xFFF00000u, 0u, 0u, 0u);
    uint32_t w[4];
    ck::philox(0u, attempt, ck::TAG_START, 1u, seed, i, w);
    const uint32_t k = kmin + ck::mod_u32(w[0], kmax - kmin + 1u);
    for (uint32_t t = 0; t < k; t++) {
      if ((t & 3u) == 0u) ck::philox(t >> 2, attempt, ck::TAG_START, 0u, seed, i, w);
      uint32_t ld;
      if (ck::ply_random<false>(L, w[t & 3u], &ld).over) break;
    }
    done = !reject || !ck::game_over(L);
  }
  if (!done) atomicOr(&work->overflow, 1u);
  uint32_t p1, p2, kings;
  ck::lane_to_abs(L, p1, p2, kings);
  mi_pos q;
  q.p1 = p1;
  q.p2 = p2;
  q.kings = kings;
  q.turn = (uint8_t)L.turn;
  q.counter = (uint8_t)(L.counter > 255u ? 255u : L.counter);
  q.pad[0] = 0;
  q.pad[1] = 0;
  out[i] = q;
}

__global__ void __launch_bounds__(256) k_divmod(const uint32_t *__restrict__ r,
                                                const uint32_t *__restrict__ n, uint32_t count,
                                                uint32_t *__restrict__ q, uint32_t *__restrict__ rem) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t qq, rr;
  ck::divmod_u32(r[i], n[i], qq, rr);
  q[i] = qq;
  // the remainder, or all ones when the general or (below 2^18) the per-ply remainder differs
  const bool small_ok = n[i] >= (1u << 18) || ck::mod_small(r[i], n[i]) == rr;
  rem[i] = ck::mod_u32(r[i], n[i]) == rr && small_ok ? rr : 0xFFFFFFFFu;
}

// Plain grid-stride copy of 16-byte words (traffic-counter calibration).
__global__ void __launch_bounds__(256) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                              uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
}

// Integer VALU issue-rate probe (the roofline's peak, measured).  Every lane runs eight
// independent chains of v_add_u32 / v_xor_b32 with uniform (SGPR) operands, written as asm so
// that the compiler can neither fuse them (v_xad_u32, v_bitop3_b32) nor fold them: one loop trip
// is exactly 32 VALU instructions, so lane-ops per launch = threads x iters x 32.
#define VALU_STEP(a)                                                                     \
  asm volatile("v_add_u32 %0, %1, %0\n\tv_xor_b32 %0, %2, %0" : "+v"(a) : "s"(k0), "s"(k1))

__global__ void __launch_bounds__(256) k_valu_rate(uint32_t iters, uint32_t salt, uint32_t *out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k0 = salt | 1u, k1 = salt * 0x9E3779B9u;
  uint32_t a0 = t ^ salt, a1 = a0 + 1u, a2 = a0 + 2u, a3 = a0 + 3u;
  uint32_t a4 = a0 + 4u, a5 = a0 + 5u, a6 = a0 + 6u, a7 = a0 + 7u;
  for (uint32_t i = 0; i < iters; i++) {
    VALU_STEP(a0); VALU_STEP(a1); VALU_STEP(a2); VALU_STEP(a3);
    VALU_STEP(a4); VALU_STEP(a5); VALU_STEP(a6); VALU_STEP(a7);
    VALU_STEP(a0); VALU_STEP(a1); VALU_STEP(a2); VALU_STEP(a3);
    VALU_STEP(a4); VALU_STEP(a5); VALU_STEP(a6); VALU_STEP(a7);
  }
  out[t] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}
#undef VALU_STEP

template <uint32_t POLICY, bool TRACE>
hipError_t launch_refill(const PlayoutArgs &args, uint32_t refill, uint32_t grid, uint32_t block,
                         hipStream_t stream) {
  if (refill == MI_REFILL_STATIC)
    hipLaunchKernelGGL((k_playouts<POLICY, TRACE, MI_REFILL_STATIC>), dim3(grid), dim3(block), 0,
                       stream, args);
  else
    hipLaunchKernelGGL((k_playouts<POLICY, TRACE, MI_REFILL_QUEUE>), dim3(grid), dim3(block), 0,
                       stream, args);
  return hipGetLastError();
}

}  // namespace

const void *playout_kernel(uint32_t policy, uint32_t refill) {
  if (policy == MI_POLICY_RANDOM)
    return refill == MI_REFILL_STATIC
               ? (const void *)k_playouts<MI_POLICY_RANDOM, false, MI_REFILL_STATIC>
               : (const void *)k_playouts<MI_POLICY_RANDOM, false, MI_REFILL_QUEUE>;
  if (policy == MI_POLICY_HEURISTIC)
    return refill == MI_REFILL_STATIC
               ? (const void *)k_playouts<MI_POLICY_HEURISTIC, false, MI_REFILL_STATIC>
               : (const void *)k_playouts<MI_POLICY_HEURISTIC, false, MI_REFILL_QUEUE>;
  return refill == MI_REFILL_STATIC
             ? (const void *)k_playouts<MI_POLICY_HEURISTIC_NOISY, false, MI_REFILL_STATIC>
             : (const void *)k_playouts<MI_POLICY_HEURISTIC_NOISY, false, MI_REFILL_QUEUE>;
}

const void *reset_kernel() { return (const void *)k_reset; }

hipError_t launch_reset(WorkWords *work, hipStream_t stream) {
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, work);
  return hipGetLastError();
}

hipError_t launch_playouts(const PlayoutArgs &args, uint32_t policy, bool trace, uint32_t refill,
                           uint32_t grid, uint32_t block, hipStream_t stream) {
  if (policy == MI_POLICY_RANDOM)
    return trace ? launch_refill<MI_POLICY_RANDOM, true>(args, refill, grid, block, stream)
                 : launch_refill<MI_POLICY_RANDOM, false>(args, refill, grid, block, stream);
  if (policy == MI_POLICY_HEURISTIC)
    return trace ? launch_refill<MI_POLICY_HEURISTIC, true>(args, refill, grid, block, stream)
                 : launch_refill<MI_POLICY_HEURISTIC, false>(args, refill, grid, block, stream);
  return trace ? launch_refill<MI_POLICY_HEURISTIC_NOISY, true>(args, refill, grid, block, stream)
               : launch_refill<MI_POLICY_HEURISTIC_NOISY, false>(args, refill, grid, block, stream);
}

hipError_t launch_gen_moves(const mi_pos *pos, uint32_t npos, mi_move *moves, uint32_t cap,
                            uint32_t *counts, WorkWords *work, hipStream_t stream) {
  const uint32_t block = 256;
  const uint32_t grid = (npos + block - 1) / block;
  hipLaunchKernelGGL(k_gen_moves, dim3(grid), dim3(block), 0, stream, pos, npos, moves, cap,
                     counts, work);
  return hipGetLastError();
}

hipError_t launch_start_positions(uint32_t seed, uint32_t n, uint32_t kmin, uint32_t kmax,
                                  uint32_t reject_terminal, mi_pos *out, WorkWords *work,
                                  hipStream_t stream) {
  const uint32_t block = 256;
  const uint32_t grid = (n + block - 1) / block;
  hipLaunchKernelGGL(k_start_positions, dim3(grid), dim3(block), 0, stream, seed, n, kmin, kmax,
                     reject_terminal, out, work);
  return hipGetLastError();
}

hipError_t launch_divmod(const uint32_t *r, const uint32_t *n, uint32_t count, uint32_t *q,
                         uint32_t *rem, hipStream_t stream) {
  const uint32_t block = 256;
  const uint32_t grid = (count + block - 1) / block;
  hipLaunchKernelGGL(k_divmod, dim3(grid), dim3(block), 0, stream, r, n, count, q, rem);
  return hipGetLastError();
}

hipError_t launch_copy(const uint4 *src, uint4 *dst, uint64_t n16, hipStream_t stream) {
  const uint32_t block = 256;
  const uint64_t want = (n16 + block - 1) / block;
  const uint32_t grid = (uint32_t)(want < 16384u ? want : 16384u);
  hipLaunchKernelGGL(k_copy, dim3(grid ? grid : 1u), dim3(block), 0, stream, src, dst, n16);
  return hipGetLastError();
}

hipError_t launch_valu_rate(uint32_t iters, uint32_t salt, uint32_t *out, uint32_t threads,
                            hipStream_t stream) {
  const uint32_t block = 256;
  hipLaunchKernelGGL(k_valu_rate, dim3(threads / block), dim3(block), 0, stream, iters, salt, out);
  return hipGetLastError();
}

hipError_t launch_tally(const int8_t *winners, const uint32_t *offs, uint32_t npos,
                        uint32_t per_pos, uint32_t *tallies, hipStream_t stream) {
  const uint32_t block = 256;
  const uint64_t threads = (uint64_t)npos * 64u;
  const uint32_t grid = (uint32_t)((threads + block - 1) / block);
  hipLaunchKernelGGL(k_tally, dim3(grid), dim3(block), 0, stream, winners, offs, npos, per_pos,
                     tallies);
  return hipGetLastError();
}

}  // namespace gpuai
