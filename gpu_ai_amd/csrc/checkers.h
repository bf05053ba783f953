// checkers.h -- 8x8 checkers on 32-bit bitboards, one position per lane.
//
// The game rules are those of gpu_ai's host path (src/state.cu, src/state.cpp,
// src/heuristic.cu); the representation is this project's own.  A position is held in the
// frame of the side to move ("mover frame"): `own` are the mover's pieces, `opp` the
// opponent's, and the mover's men advance toward higher rows.  Player 1's absolute frame is the
// mover frame; player 2's is the 180-degree rotation, which maps square s to 31 - s and is a
// single bit reversal.
//
// Reference move order and the mover frame.  State::genMoves lists moves square by square in
// index order (src/state.cu:171-180).  Within a square, direct moves go (+1,+1), (+1,-1),
// (-1,+1), (-1,-1) restricted to the piece's directions (src/state.cu:255-281); capture
// sequences are the leaves of a depth-first search that tries a man's (f,-1) jump before its
// (f,+1) jump (src/state.cu:283-340) and a king's four jumps in the direction order above
// (src/state.cu:388-420).  Rotating the board by 180 degrees reverses the square order and
// reverses every per-square and per-node order, so for player 2 the list enumerated in the
// mover frame with player 1's orders is exactly the reference list reversed: index k of the
// reference list is index n-1-k here.
//
// Square geometry: s = 4*row + i, i = s & 3; col = 2i + 1 on even rows and 2i on odd rows.
// One-step neighbours differ by +5/+4 (even/odd row) for (+1,+1), +4/+3 for (+1,-1), -3/-4 for
// (-1,+1) and -4/-5 for (-1,-1); two-step jumps by +9, +7, -7, -9 on every row.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CK_HD __host__ __device__ __forceinline__

namespace ck {

// ------------------------------------------------------------------------------------------
// bit helpers
// ------------------------------------------------------------------------------------------
CK_HD uint32_t popc(uint32_t x) { return (uint32_t)__builtin_popcount(x); }
CK_HD uint32_t ctz(uint32_t x) { return (uint32_t)__builtin_ctz(x); }
CK_HD uint32_t brev(uint32_t x) { return __builtin_bitreverse32(x); }
CK_HD uint32_t bit(uint32_t s) { return 1u << s; }
CK_HD uint32_t getbit(uint32_t m, uint32_t s) { return (m >> s) & 1u; }

// ------------------------------------------------------------------------------------------
// geometry
// ------------------------------------------------------------------------------------------
constexpr uint32_t EVEN_ROWS = 0x0F0F0F0Fu;  // rows 0, 2, 4, 6
constexpr uint32_t ODD_ROWS = 0xF0F0F0F0u;   // rows 1, 3, 5, 7
constexpr uint32_t I0 = 0x11111111u;         // i == 0
constexpr uint32_t I3 = 0x88888888u;         // i == 3
constexpr uint32_t ROW5 = 0x00F00000u;       // a man's jump from here lands on the far row
constexpr uint32_t ROW6 = 0x0F000000u;
constexpr uint32_t ROW7 = 0xF0000000u;       // the mover's far row (promotion)
constexpr uint32_t FAR_ROW_FIRST = 28u;

// Squares reached by one step from the set b, per direction (off-board steps vanish).
CK_HD uint32_t step0(uint32_t b) { return ((b & (EVEN_ROWS & ~I3)) << 5) | ((b & ODD_ROWS) << 4); }
CK_HD uint32_t step1(uint32_t b) { return ((b & EVEN_ROWS) << 4) | ((b & (ODD_ROWS & ~I0)) << 3); }
CK_HD uint32_t step2(uint32_t b) { return ((b & (EVEN_ROWS & ~I3)) >> 3) | ((b & ODD_ROWS) >> 4); }
CK_HD uint32_t step3(uint32_t b) { return ((b & EVEN_ROWS) >> 4) | ((b & (ODD_ROWS & ~I0)) >> 5); }

// Jump displacement of direction d: (+1,+1) +9, (+1,-1) +7, (-1,+1) -7, (-1,-1) -9.
CK_HD int jump_off(uint32_t d) {
  int j = ((d ^ (d >> 1)) & 1u) ? 7 : 9;
  return (d & 2u) ? -j : j;
}
// Square one step from x in direction d (the step is half the jump, rounded by row parity).
CK_HD uint32_t step_to(uint32_t x, uint32_t d) {
  int j = jump_off(d);
  int odd = (int)((x >> 2) & 1u);
  return (uint32_t)((int)x + ((j + 1 - 2 * odd) >> 1));
}
// Square jumped when moving from x to land (two rows apart).
CK_HD uint32_t mid_sq(uint32_t x, uint32_t land) { return ((x + land + 1u) >> 1) - ((x >> 2) & 1u); }

// ------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// ------------------------------------------------------------------------------------------
CK_HD void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                  uint32_t out[4]) {
#pragma unroll
  for (int round = 0; round < 10; round++) {
    uint64_t a = (uint64_t)0xD2511F53u * c0;
    uint64_t b = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)b;
    c3 = (uint32_t)a;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

constexpr uint32_t TAG_PLAYOUT = 0u;
constexpr uint32_t TAG_START = 1u;
constexpr uint32_t TAG_SEARCH = 2u;
constexpr uint32_t TAG_MLP = 3u;

// Reciprocal of n in double precision: a seed (on the device the hardware estimate v_rcp_f64, of
// roughly single precision; on the host 1/n) refined by two Newton-Raphson steps, each of which
// squares the relative error, so the result is within a few units of 2^-52 of 1/n.  The
// remainder routines below run the same code on the host and the device, so the host's checks
// of the lane logic exercise them on every ply.
CK_HD double recip(double dn) {
#if defined(__HIP_DEVICE_COMPILE__)
  double x = __builtin_amdgcn_rcp(dn);
#else
  double x = 1.0 / dn;
#endif
  x = __builtin_fma(x, __builtin_fma(-dn, x, 1.0), x);
  x = __builtin_fma(x, __builtin_fma(-dn, x, 1.0), x);
  return x;
}

// Exact r mod n for 1 <= n: the refined reciprocal, a floor, an exact fused remainder, and one
// correction each way (r / n is estimated to far better than one unit, so the floor is off by at
// most one).  The remainder r - q*n is an integer in (-n, 2n), exact in double precision, and is
// corrected there: as a 32-bit integer it would overflow for n >= 2^31.
CK_HD uint32_t mod_u32(uint32_t r, uint32_t n) {
  const double dn = (double)n, dr = (double)r;
  const double q = __builtin_floor(dr * recip(dn));
  double m = __builtin_fma(-q, dn, dr);
  m = m < 0.0 ? m + dn : (m >= dn ? m - dn : m);
  return (uint32_t)m;
}

// Exact r mod n for 1 <= n < 2^18 (move and tie counts, the per-ply draw), with no correction
// step.  After two Newton steps from a seed of at least 13 bits the reciprocal x has a relative
// error below 2^-52, so |r*x - r/n| plus the rounding of the fused sum stays under
// (2^32 / n)(2^-52 + 2^-53) = 1.5 * 2^-20 / n.  Adding 2^-19 lifts fma(r, x, 2^-19) to at least
// r/n; it stays below floor(r/n) + 1 because frac(r/n) <= 1 - 1/n and
// 1.5 * 2^-20 / n + 2^-19 < 1/n while n < 2^18.  The floor is then the exact quotient, and the
// fused remainder r - q*n is exact.
CK_HD uint32_t mod_small(uint32_t r, uint32_t n) {
  const double dn = (double)n, dr = (double)r;
  const double q = __builtin_floor(__builtin_fma(dr, recip(dn), 0x1p-19));
  return (uint32_t)__builtin_fma(-q, dn, dr);
}

// Exact quotient and remainder of r / n for 1 <= n (same estimate-and-correct scheme, the
// correction in double precision).
CK_HD void divmod_u32(uint32_t r, uint32_t n, uint32_t &q, uint32_t &rem) {
  const double dn = (double)n, dr = (double)r;
  double qd = __builtin_floor(dr * recip(dn));
  double m = __builtin_fma(-qd, dn, dr);
  if (m < 0.0) {
    m += dn;
    qd -= 1.0;
  } else if (m >= dn) {
    m -= dn;
    qd += 1.0;
  }
  q = (uint32_t)qd;
  rem = (uint32_t)m;
}

// ------------------------------------------------------------------------------------------
// positions
// ------------------------------------------------------------------------------------------
struct Lane {
  uint32_t own, opp, kings;  // mover frame
  uint32_t turn;             // absolute side to move: 0 player 1, 1 player 2
  uint32_t counter;          // plies since the last capture
};

CK_HD Lane lane_from_abs(uint32_t p1, uint32_t p2, uint32_t kings, uint32_t turn,
                         uint32_t counter) {
  Lane L;
  if (turn == 0) {
    L.own = p1;
    L.opp = p2;
    L.kings = kings & (p1 | p2);
  } else {
    L.own = brev(p2);
    L.opp = brev(p1);
    L.kings = brev(kings & (p1 | p2));
  }
  L.turn = turn;
  L.counter = counter;
  return L;
}

CK_HD void lane_to_abs(const Lane &L, uint32_t &p1, uint32_t &p2, uint32_t &kings) {
  if (L.turn == 0) {
    p1 = L.own;
    p2 = L.opp;
    kings = L.kings;
  } else {
    p1 = brev(L.opp);
    p2 = brev(L.own);
    kings = brev(L.kings);
  }
}

// Absolute square of a mover-frame square.
CK_HD uint32_t abs_sq(const Lane &L, uint32_t s) { return L.turn ? 31u - s : s; }

// ------------------------------------------------------------------------------------------
// move-list analysis of one position
// ------------------------------------------------------------------------------------------
// cap[d]  : squares from which a jump in direction d is valid on the unmodified board (an
//           opponent piece one step away, an empty square two steps away; src/state.cu:122-132)
// slot[0..3]: per-square move sources in reference order, for every square that has no
//           multi-jump: captures (men: (+1,-1) then (+1,+1); kings: directions 0..3) when
//           captures exist, else direct moves (directions 0..3; men only 0 and 1)
// multi   : capture origins with a sequence longer than one jump (need the tree walk)
struct Gen {
  uint32_t cap[4];
  uint32_t slot[4];
  uint32_t multi;
  uint32_t capture;  // captures are mandatory this ply (src/state.cu:239-245)
  uint32_t origins;  // capture origins (all, in square order)
};

CK_HD void analyse(const Lane &L, Gen &g) {
  const uint32_t empty = ~(L.own | L.opp);
  const uint32_t men = L.own & ~L.kings;
  const uint32_t kng = L.own & L.kings;
  g.cap[0] = step3(L.opp & step3(empty));
  g.cap[1] = step2(L.opp & step2(empty));
  g.cap[2] = step1(L.opp & step1(empty));
  g.cap[3] = step0(L.opp & step0(empty));
  const uint32_t fwd = g.cap[0] | g.cap[1];
  const uint32_t any = fwd | g.cap[2] | g.cap[3];
  const uint32_t csA = (men & g.cap[1]) | (kng & g.cap[0]);
  const uint32_t csB = (men & g.cap[0]) | (kng & g.cap[1]);
  const uint32_t csC = kng & g.cap[2];
  const uint32_t csD = kng & g.cap[3];
  g.origins = csA | csB | csC | csD;
  g.capture = g.origins != 0u;
  // An origin needs the walk if a first jump lands on a square with a further valid jump.  For
  // a king the jump straight back would land on its own (occupied) origin, so it never counts.
  g.multi = (men & ((g.cap[1] & (fwd >> 7)) | (g.cap[0] & (fwd >> 9)))) |
            (kng & ((g.cap[0] & (any >> 9)) | (g.cap[1] & (any >> 7)) | (g.cap[2] & (any << 7)) |
                    (g.cap[3] & (any << 9))));
  if (g.capture) {
    g.slot[0] = csA;
    g.slot[1] = csB;
    g.slot[2] = csC;
    g.slot[3] = csD;
  } else {
    g.slot[0] = L.own & step3(empty);
    g.slot[1] = L.own & step2(empty);
    g.slot[2] = kng & step1(empty);
    g.slot[3] = kng & step0(empty);
  }
}

CK_HD uint32_t slot_count(const uint32_t s[4]) { return popc(s[0]) + popc(s[1]) + popc(s[2]) + popc(s[3]); }

// Direction of slot `sl` for the piece on square `sq`.
CK_HD uint32_t slot_dir(const Lane &L, const Gen &g, uint32_t sq, uint32_t sl) {
  uint32_t king = getbit(L.kings, sq);
  return (g.capture && !king && sl < 2u) ? (sl ^ 1u) : sl;
}

// Byte permutation (v_perm_b32): byte i of the result is byte sel_i of the 8-byte value
// {hi:lo} (0-3 from lo, 4-7 from hi), or 0 for a selector of 12.
CK_HD uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_perm(hi, lo, sel);
#else
  const uint64_t v = ((uint64_t)hi << 32) | lo;
  uint32_t r = 0;
  for (int i = 0; i < 4; i++) {
    const uint32_t b = (sel >> (8 * i)) & 0xFFu;
    const uint32_t byte = b < 8u ? (uint32_t)(v >> (8u * b)) & 0xFFu : 0u;
    r |= byte << (8 * i);
  }
  return r;
#endif
}

// k-th move (0-based) of the slot masks in square-major, slot-minor order.  The four masks are
// transposed by bytes, so that word x_b holds squares 8b..8b+7 of slot t in its byte t; the
// group of eight squares holding the k-th move is found from the four words' popcounts, and the
// square inside it by a 3-step search that counts the group's moves below a square with one
// popcount (the square's low-bit mask copied into every byte).
CK_HD void select_slot(const uint32_t s[4], uint32_t k, uint32_t &sq, uint32_t &sl) {
  const uint32_t a_lo = perm(s[1], s[0], 0x05010400u), a_hi = perm(s[1], s[0], 0x07030602u);
  const uint32_t b_lo = perm(s[3], s[2], 0x05010400u), b_hi = perm(s[3], s[2], 0x07030602u);
  const uint32_t x0 = perm(b_lo, a_lo, 0x05040100u), x1 = perm(b_lo, a_lo, 0x07060302u);
  const uint32_t x2 = perm(b_hi, a_hi, 0x05040100u), x3 = perm(b_hi, a_hi, 0x07060302u);
  const uint32_t p1 = popc(x0), p2 = p1 + popc(x1), p3 = p2 + popc(x2);
  const bool g1 = k >= p1, g2 = k >= p2, g3 = k >= p3;
  const uint32_t x = g3 ? x3 : g2 ? x2 : g1 ? x1 : x0;
  const uint32_t base = g3 ? p3 : g2 ? p2 : g1 ? p1 : 0u;
  const uint32_t kk = k - base;
  auto below = [x](uint32_t j) { return popc(x & perm(0u, bit(j) - 1u, 0u)); };
  uint32_t lo = 0;
#pragma unroll
  for (uint32_t stride = 4; stride >= 1; stride >>= 1)
    if (below(lo + stride) <= kk) lo += stride;
  uint32_t j = kk - below(lo);
  const uint32_t y = (x >> lo) & 0x01010101u;
  uint32_t here = (y | (y >> 7) | (y >> 14) | (y >> 21)) & 15u;
  if (j >= 1u) here &= here - 1u;
  if (j >= 2u) here &= here - 1u;
  if (j >= 3u) here &= here - 1u;
  sq = 8u * (uint32_t)(g1 + g2 + g3) + lo;
  sl = ctz(here);
}

// ------------------------------------------------------------------------------------------
// capture-tree walk (src/state.cu:283-340 for men, 388-420 for kings)
// ------------------------------------------------------------------------------------------
// Children of the node on square x, as a bit mask in reference order.  Men: bit 0 = (+1,-1),
// bit 1 = (+1,+1).  Kings: bit d = direction d, minus landings already used by the sequence
// (the cycle check of src/state.cu:134-139).
CK_HD uint32_t kids(uint32_t x, uint32_t king, const uint32_t cap[4], uint32_t visited) {
  if (!king) return getbit(cap[1], x) | (getbit(cap[0], x) << 1);
  uint32_t c0 = cap[0] & ~(visited >> 9);
  uint32_t c1 = cap[1] & ~(visited >> 7);
  uint32_t c2 = cap[2] & ~(visited << 7);
  uint32_t c3 = cap[3] & ~(visited << 9);
  return getbit(c0, x) | (getbit(c1, x) << 1) | (getbit(c2, x) << 2) | (getbit(c3, x) << 3);
}

// Depth-first walk of the capture sequences of the piece on `o`.  For each leaf, in order,
// calls visit(to, removed_mask, jumps, dirs) where dirs holds the direction of jump t in bits
// 2t..2t+1; the walk stops when visit returns true.  A king sequence has at most 7 jumps: all
// its landing squares share one class of 8 squares ((row mod 2, (row+col) mod 4) is invariant
// under a jump) with the origin, which stays occupied; men make at most 3.
template <class Visit>
CK_HD void walk(uint32_t o, uint32_t king, const uint32_t cap[4], Visit &visit) {
  uint32_t x = o, depth = 0, visited = 0, removed = 0, dirs = 0;
  uint32_t rem = kids(o, king, cap, 0);
  for (;;) {
    uint32_t cur = (rem >> (4u * depth)) & 15u;
    if (cur) {
      uint32_t s = ctz(cur);
      rem ^= bit(4u * depth + s);
      uint32_t d = king ? s : (s ^ 1u);
      uint32_t land = (uint32_t)((int)x + jump_off(d));
      visited |= bit(land);
      removed |= bit(mid_sq(x, land));
      dirs = (dirs & ~(3u << (2u * depth))) | (d << (2u * depth));
      depth++;
      x = land;
      uint32_t ch = depth < 7u ? kids(x, king, cap, visited) : 0u;
      if (ch) {
        rem = (rem & ~(15u << (4u * depth))) | (ch << (4u * depth));
        continue;
      }
      if (visit(x, removed, depth, dirs)) return;
    } else if (depth == 0) {
      return;
    }
    depth--;
    uint32_t d = (dirs >> (2u * depth)) & 3u;
    uint32_t prev = (uint32_t)((int)x - jump_off(d));
    visited &= ~bit(x);
    removed &= ~bit(mid_sq(prev, x));
    x = prev;
  }
}

// ------------------------------------------------------------------------------------------
// move application (src/state.cu:57-92), then hand the turn over (rotate into the new mover's
// frame)
// ------------------------------------------------------------------------------------------
CK_HD void apply(Lane &L, uint32_t from, uint32_t to, uint32_t removed) {
  const uint32_t fb = bit(from), tb = bit(to);
  const uint32_t was_king = getbit(L.kings, from);
  uint32_t own = L.own ^ (fb | tb);
  uint32_t opp = L.opp & ~removed;
  uint32_t kings = L.kings & ~removed;
  kings = was_king ? (kings ^ (fb | tb)) : (kings | ((to >= FAR_ROW_FIRST) ? tb : 0u));
  L.counter = removed ? 0u : L.counter + 1u;
  L.own = brev(opp);
  L.opp = brev(own);
  L.kings = brev(kings);
  L.turn ^= 1u;
}

// ------------------------------------------------------------------------------------------
// heuristic weights (src/heuristic.cu:7-49): material with man 1, king 4
// ------------------------------------------------------------------------------------------
CK_HD uint32_t material(uint32_t pieces, uint32_t kings) {
  return popc(pieces & ~kings) + 4u * popc(pieces & kings);
}

// ------------------------------------------------------------------------------------------
// move-list hashing (verification mode).  Record layout shared with the oracle:
//   move digest d = fmix(0x6A09E667 ^ (from | to<<5 | jumps<<10 | promoted<<14)), then per jump
//                   d = fmix(d ^ (inter_t | removed_t<<5 | (t+1)<<10))        (absolute squares)
//   list digest    = sum_i fmix(d_i + i * 0x9E3779B9)                         (reference index i)
//   ply record     = fmix(list ^ fmix(n | chosen<<16)), chosen = 0xFFFF at game end
//   chain          h' = fmix(h ^ record) + 0x9E3779B9, starting at 0
// ------------------------------------------------------------------------------------------
CK_HD uint32_t fmix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}
CK_HD uint32_t digest_head(uint32_t from, uint32_t to, uint32_t jumps, uint32_t promoted) {
  return fmix(0x6A09E667u ^ (from | (to << 5) | (jumps << 10) | (promoted << 14)));
}
CK_HD uint32_t digest_jump(uint32_t d, uint32_t inter, uint32_t removed, uint32_t t) {
  return fmix(d ^ (inter | (removed << 5) | ((t + 1u) << 10)));
}
CK_HD uint32_t ply_record(uint32_t list, uint32_t n, uint32_t chosen) {
  return fmix(list ^ fmix(n | (chosen << 16)));
}
CK_HD uint32_t hash_step(uint32_t h, uint32_t rec) { return fmix(h ^ rec) + 0x9E3779B9u; }

// Visits every move of the position in mover-frame order:
//   visit(index, from, to, removed_mask, jumps, dirs, king)
// (jumps == 0 for direct moves).  Used by verification, hashing and the host move generator.
template <class Visit>
CK_HD uint32_t enumerate(const Lane &L, const Gen &g, Visit &visit) {
  uint32_t idx = 0;
  if (g.capture) {
    uint32_t o = g.origins;
    while (o) {
      uint32_t s = ctz(o);
      o &= o - 1u;
      uint32_t king = getbit(L.kings, s);
      auto leaf = [&](uint32_t to, uint32_t removed, uint32_t jumps, uint32_t dirs) {
        visit(idx, s, to, removed, jumps, dirs, king);
        idx++;
        return false;
      };
      walk(s, king, g.cap, leaf);
    }
  } else {
    uint32_t any = g.slot[0] | g.slot[1] | g.slot[2] | g.slot[3];
    while (any) {
      uint32_t s = ctz(any);
      any &= any - 1u;
      uint32_t king = getbit(L.kings, s);
      for (uint32_t sl = 0; sl < 4u; sl++) {
        if (!getbit(g.slot[sl], s)) continue;
        visit(idx, s, step_to(s, sl), 0u, 0u, 0u, king);
        idx++;
      }
    }
  }
  return idx;
}

// Digest of one move given its mover-frame description.
CK_HD uint32_t move_digest(const Lane &L, uint32_t from, uint32_t to, uint32_t jumps,
                           uint32_t dirs, uint32_t king) {
  uint32_t promoted = (!king && to >= FAR_ROW_FIRST) ? 1u : 0u;
  uint32_t d = digest_head(abs_sq(L, from), abs_sq(L, to), jumps, promoted);
  uint32_t x = from;
  for (uint32_t t = 0; t < jumps; t++) {
    uint32_t dir = (dirs >> (2u * t)) & 3u;
    uint32_t land = (uint32_t)((int)x + jump_off(dir));
    d = digest_jump(d, abs_sq(L, land), abs_sq(L, mid_sq(x, land)), t);
    x = land;
  }
  return d;
}

// List digest with reference indices (index n-1-i for player 2).
CK_HD uint32_t list_digest(const Lane &L, const Gen &g, uint32_t n) {
  uint32_t acc = 0;
  auto v = [&](uint32_t i, uint32_t from, uint32_t to, uint32_t, uint32_t jumps, uint32_t dirs,
               uint32_t king) {
    uint32_t ref_i = L.turn ? n - 1u - i : i;
    acc += fmix(move_digest(L, from, to, jumps, dirs, king) + ref_i * 0x9E3779B9u);
  };
  enumerate(L, g, v);
  return acc;
}

// ------------------------------------------------------------------------------------------
// one ply
// ------------------------------------------------------------------------------------------
// Result of choosing: the move in mover-frame squares.
struct Choice {
  uint32_t from, to, removed;
};

// Number of capture sequences of all origins (random policy, multi-jump ply).
CK_HD uint32_t count_walks(const Lane &L, const Gen &g) {
  uint32_t n = 0;
  uint32_t o = g.origins;
  auto leaf = [&](uint32_t, uint32_t, uint32_t, uint32_t) {
    n++;
    return false;
  };
  while (o) {
    uint32_t s = ctz(o);
    o &= o - 1u;
    walk(s, getbit(L.kings, s), g.cap, leaf);
  }
  return n;
}

// k-th capture sequence in mover-frame order.
CK_HD Choice select_walk(const Lane &L, const Gen &g, uint32_t k) {
  Choice c = {0u, 0u, 0u};
  uint32_t o = g.origins;
  uint32_t left = k;
  bool found = false;
  auto leaf = [&](uint32_t to, uint32_t removed, uint32_t, uint32_t) {
    if (left == 0u) {
      c.to = to;
      c.removed = removed;
      found = true;
      return true;
    }
    left--;
    return false;
  };
  while (o && !found) {
    uint32_t s = ctz(o);
    o &= o - 1u;
    c.from = s;
    walk(s, getbit(L.kings, s), g.cap, leaf);
  }
  return c;
}

// Move of slot choice (square sq, slot sl).  A step in direction d from square x moves
// base_d - (x's row parity) squares, base = {5, 4, -3, -4}; a jump is two steps from rows of
// opposite parity, 2 base_d - 1 squares, over the square one step away.
CK_HD Choice slot_choice(const Lane &L, const Gen &g, uint32_t sq, uint32_t sl) {
  const uint32_t d = slot_dir(L, g, sq, sl);
  const int base = (int)((int32_t)(0xFCFD0405u << (24u - 8u * d)) >> 24);
  const int step = base - (int)((sq >> 2) & 1u);
  Choice c;
  c.from = sq;
  c.to = (uint32_t)((int)sq + (g.capture ? 2 * base - 1 : step));
  c.removed = g.capture ? bit((uint32_t)((int)sq + step)) : 0u;
  return c;
}

// Leaf count of the capture tree of one origin.
CK_HD uint32_t count_walk(const Lane &L, const Gen &g, uint32_t s) {
  uint32_t n = 0;
  auto leaf = [&](uint32_t, uint32_t, uint32_t, uint32_t) {
    n++;
    return false;
  };
  walk(s, getbit(L.kings, s), g.cap, leaf);
  return n;
}

// Leaf `k` (mover-frame order) of the capture tree of origin s.
CK_HD Choice select_in_walk(const Lane &L, const Gen &g, uint32_t s, uint32_t k) {
  Choice c = {s, 0u, 0u};
  uint32_t left = k;
  auto leaf = [&](uint32_t to, uint32_t removed, uint32_t, uint32_t) {
    if (left == 0u) {
      c.to = to;
      c.removed = removed;
      return true;
    }
    left--;
    return false;
  };
  walk(s, getbit(L.kings, s), g.cap, leaf);
  return c;
}

// The first four capture sequences of the multi-jump kings (in walk order): landing square and
// removed pieces, held in registers (fixed slots, selected without dynamic indexing) so that
// choosing one of them needs no second walk.
struct LeafCache {
  uint32_t to0, to1, to2, to3;
  uint32_t rm0, rm1, rm2, rm3;
};

CK_HD void cache_put(LeafCache &c, uint32_t i, uint32_t to, uint32_t rm) {
  c.to0 = i == 0u ? to : c.to0;
  c.rm0 = i == 0u ? rm : c.rm0;
  c.to1 = i == 1u ? to : c.to1;
  c.rm1 = i == 1u ? rm : c.rm1;
  c.to2 = i == 2u ? to : c.to2;
  c.rm2 = i == 2u ? rm : c.rm2;
  c.to3 = i == 3u ? to : c.to3;
  c.rm3 = i == 3u ? rm : c.rm3;
}

// Reads entry i (< 4) with masks rather than a select chain: a chain of selects over the fields
// is folded by the compiler into one load through a selected address, which puts the whole
// cache in scratch memory.
CK_HD void cache_get(const LeafCache &c, uint32_t i, uint32_t &to, uint32_t &rm) {
  const uint32_t m0 = 0u - (uint32_t)(i == 0u), m1 = 0u - (uint32_t)(i == 1u);
  const uint32_t m2 = 0u - (uint32_t)(i == 2u), m3 = 0u - (uint32_t)(i == 3u);
  to = (c.to0 & m0) | (c.to1 & m1) | (c.to2 & m2) | (c.to3 & m3);
  rm = (c.rm0 & m0) | (c.rm1 & m1) | (c.rm2 & m2) | (c.rm3 & m3);
}

// ------------------------------------------------------------------------------------------
// a man's capture tree in closed form
// ------------------------------------------------------------------------------------------
// A man only jumps forward (+7 first in the reference's order, then +9, src/state.cu:292-302),
// and every jump's validity is read on the unmodified board, so whether a jump from a square is
// valid does not depend on the path that reached it; with two rows gained per jump a sequence
// has at most three jumps.  The leaves of the tree therefore fall on eight fixed slots, one per
// direction triple (d1, d2, d3) (d = 0 for +7, 1 for +9), slot 4*d1 + 2*d2 + d3; a sequence that
// ends after one or two jumps takes the first slot of its subtree (trailing directions 0).  The
// DFS order of the leaves is the slot order, so the tree is an 8-bit mask of present slots.
// Bit r of `a` / `b` (the cap masks shifted right by the origin) says whether a +7 / +9 jump is
// valid from the square r above the origin; squares past the board read as 0.
CK_HD uint32_t man_tree(const Gen &g, uint32_t o) {
  const uint32_t a = g.cap[1] >> o, b = g.cap[0] >> o, h = a | b;
  auto at = [](uint32_t m, uint32_t r) { return (m >> r) & 1u; };
  const uint32_t a0 = at(a, 0), b0 = at(b, 0);
  const uint32_t h7 = at(h, 7), h9 = at(h, 9), h14 = at(h, 14), h16 = at(h, 16), h18 = at(h, 18);
  // subtree of the landing square 7 (first jump +7) and of 9 (+9); a landing square with no
  // further jump is a leaf in the subtree's first slot
  const uint32_t t7 = a0 & h7, t9 = b0 & h9;
  const uint32_t s00 = t7 & at(a, 7), s01 = t7 & at(b, 7);  // 7 -> 14, 7 -> 16
  const uint32_t s10 = t9 & at(a, 9), s11 = t9 & at(b, 9);  // 9 -> 16, 9 -> 18
  uint32_t p = 0;
  p |= ((a0 & (h7 ^ 1u)) | (s00 & ((h14 ^ 1u) | at(a, 14)))) << 0;
  p |= (s00 & h14 & at(b, 14)) << 1;
  p |= (s01 & ((h16 ^ 1u) | at(a, 16))) << 2;
  p |= (s01 & h16 & at(b, 16)) << 3;
  p |= ((b0 & (h9 ^ 1u)) | (s10 & ((h16 ^ 1u) | at(a, 16)))) << 4;
  p |= (s10 & h16 & at(b, 16)) << 5;
  p |= (s11 & ((h18 ^ 1u) | at(a, 18))) << 6;
  p |= (s11 & h18 & at(b, 18)) << 7;
  return p;
}

// Leaf k (0-based, DFS order) of the tree man_tree gave for origin o: the k-th set bit of p
// names the slot, and the tree's branching says where the sequence stops.
CK_HD Choice man_leaf(const Gen &g, uint32_t o, uint32_t p, uint32_t k) {
  // k-th set bit of an 8-bit mask
  uint32_t slot = 0;
  const uint32_t c4 = popc(p & 15u);
  if (k >= c4) {
    k -= c4;
    slot = 4;
  }
  const uint32_t c2 = popc((p >> slot) & 3u);
  if (k >= c2) {
    k -= c2;
    slot += 2;
  }
  const uint32_t c1 = (p >> slot) & 1u;
  if (k >= c1) slot += 1;
  const uint32_t d1 = slot >> 2, d2 = (slot >> 1) & 1u, d3 = slot & 1u;
  const uint32_t h = g.cap[0] | g.cap[1];
  Choice c;
  c.from = o;
  uint32_t x = o + 7u + 2u * d1;
  uint32_t removed = bit(mid_sq(o, x));
  if (getbit(h, x)) {  // a second jump (a landing square with a jump never ends a sequence)
    const uint32_t y = x + 7u + 2u * d2;
    removed |= bit(mid_sq(x, y));
    x = y;
    if (getbit(h, x)) {
      const uint32_t z = x + 7u + 2u * d3;
      removed |= bit(mid_sq(x, z));
      x = z;
    }
  }
  c.to = x;
  c.removed = removed;
  return c;
}

// ------------------------------------------------------------------------------------------
// a king's capture tree to depth two in closed form
// ------------------------------------------------------------------------------------------
// A king jumps in directions 0..3 (+9, +7, -7, -9; src/state.cu:411-417) and may not land on a
// square it landed on before (src/state.cu:134-139).  Up to two jumps that rule never binds: the
// only way back is the straight jump onto the origin, which is occupied.  The tree to depth two
// therefore falls on sixteen slots 4*d1 + d2 (a sequence of one jump takes slot 4*d1), in DFS
// order.  `deep` is set when some second landing square allows a third jump other than the one
// straight back (which the rule forbids); such trees are left to the walk.
CK_HD uint32_t king_tree(const Gen &g, uint32_t o, uint32_t &deep) {
  const uint32_t any = g.cap[0] | g.cap[1] | g.cap[2] | g.cap[3];
  // third jumps from a second landing square reached in direction d (not straight back), seen
  // from the first landing square: squares x with a valid jump in d to a square that continues
  uint32_t d3 = 0;
#pragma unroll
  for (uint32_t d = 0; d < 4u; d++) {
    // a jump in any direction but 3 - d, the one straight back
    const uint32_t cont = (d == 3u ? 0u : g.cap[0]) | (d == 2u ? 0u : g.cap[1]) |
                          (d == 1u ? 0u : g.cap[2]) | (d == 0u ? 0u : g.cap[3]);
    const int off = jump_off(d);
    d3 |= g.cap[d] & (off > 0 ? cont >> off : cont << -off);
  }
  // windows around the origin: bit 32 + r of w(m) is bit o + r of m (0 off the board)
  auto w = [o](uint32_t m) { return ((uint64_t)m << 32) >> o; };
  auto at = [](uint64_t x, int r) { return (uint32_t)(x >> (32 + r)) & 1u; };
  const uint64_t w0 = w(g.cap[0]), w1 = w(g.cap[1]), w2 = w(g.cap[2]), w3 = w(g.cap[3]);
  const uint64_t wa = w(any), wd = w(d3);
  uint32_t t = 0, dp = 0;
#pragma unroll
  for (uint32_t d1 = 0; d1 < 4u; d1++) {
    const int r1 = jump_off(d1);
    const uint32_t v1 = at(d1 == 0 ? w0 : d1 == 1 ? w1 : d1 == 2 ? w2 : w3, 0);
    const uint32_t c1 = at(wa, r1);  // the jump straight back reads as invalid: origin occupied
    t |= (v1 & (c1 ^ 1u)) << (4u * d1);
    const uint32_t go = v1 & c1;
    t |= ((go & at(w0, r1)) | ((go & at(w1, r1)) << 1) | ((go & at(w2, r1)) << 2) |
          ((go & at(w3, r1)) << 3)) << (4u * d1);
    dp |= go & at(wd, r1);
  }
  deep = dp;
  return t;
}

// Leaf k (0-based, DFS order) of the tree king_tree gave for origin o.
CK_HD Choice king_leaf(const Gen &g, uint32_t o, uint32_t t, uint32_t k) {
  uint32_t slot = 0;
  const uint32_t c8 = popc(t & 255u);
  if (k >= c8) {
    k -= c8;
    slot = 8;
  }
  const uint32_t c4 = popc((t >> slot) & 15u);
  if (k >= c4) {
    k -= c4;
    slot += 4;
  }
  const uint32_t c2 = popc((t >> slot) & 3u);
  if (k >= c2) {
    k -= c2;
    slot += 2;
  }
  if (k >= ((t >> slot) & 1u)) slot += 1;
  const uint32_t any = g.cap[0] | g.cap[1] | g.cap[2] | g.cap[3];
  Choice c;
  c.from = o;
  const uint32_t x = (uint32_t)((int)o + jump_off(slot >> 2));
  uint32_t removed = bit(mid_sq(o, x));
  c.to = x;
  if (getbit(any, x)) {
    const uint32_t y = (uint32_t)((int)x + jump_off(slot & 3u));
    removed |= bit(mid_sq(x, y));
    c.to = y;
  }
  c.removed = removed;
  return c;
}

// A capture ply's moves as capture trees, origin by origin in square order: men's from
// man_tree, kings' from king_tree.  Up to four origins' masks are kept (16 bits each, two per
// word); `deep` is set when a king's tree goes past two jumps or more than four origins can
// capture (then the walk decides the ply).  Returns the move count.
CK_HD uint32_t tree_count(const Lane &L, const Gen &g, uint32_t &m01, uint32_t &m23,
                          uint32_t &deep) {
  uint32_t o = g.origins, n = 0, i = 0;
  m01 = 0;
  m23 = 0;
  deep = popc(o) > 4u;
  while (o && !deep) {
    const uint32_t s = ctz(o);
    o &= o - 1u;
    uint32_t dp = 0;
    const uint32_t t = getbit(L.kings, s) ? king_tree(g, s, dp) : man_tree(g, s);
    deep |= dp;
    if (i < 2u) m01 |= t << (16u * i);
    else m23 |= t << (16u * (i - 2u));
    n += popc(t);
    i++;
  }
  return n;
}

// k-th move (mover-frame order) from the masks of tree_count.
CK_HD Choice tree_select(const Lane &L, const Gen &g, uint32_t k, uint32_t m01, uint32_t m23) {
  uint32_t o = g.origins, i = 0;
  Choice c = {0u, 0u, 0u};
  while (o) {
    const uint32_t s = ctz(o);
    o &= o - 1u;
    const uint32_t t = ((i < 2u ? m01 : m23) >> (16u * (i & 1u))) & 0xFFFFu;
    const uint32_t n = popc(t);
    if (k < n) {
      c = getbit(L.kings, s) ? king_leaf(g, s, t, k) : man_leaf(g, s, t, k);
      break;
    }
    k -= n;
    i++;
  }
  return c;
}

// Move count of a capture ply with multi-jump origins: single-jump origins contribute their slot
// bits, each multi-jump origin its tree's leaf count (men: man_tree; kings: the walk).  The first
// four such counts are returned packed one per byte (in square order) for the selection pass,
// with the men's tree masks likewise in `trees` and the first four king leaves in `lc`.
CK_HD uint32_t count_mixed(const Lane &L, const Gen &g, uint32_t &packed, uint32_t &trees,
                           LeafCache &lc) {
  uint32_t n = popc(g.slot[0] & ~g.multi) + popc(g.slot[1] & ~g.multi) +
               popc(g.slot[2] & ~g.multi) + popc(g.slot[3] & ~g.multi);
  uint32_t o = g.multi, i = 0, m = 0;
  packed = 0;
  trees = 0;
  while (o) {
    const uint32_t s = ctz(o);
    o &= o - 1u;
    uint32_t c = 0;
    if (getbit(L.kings, s)) {
      auto leaf = [&](uint32_t to, uint32_t removed, uint32_t, uint32_t) {
        if (m < 4u) cache_put(lc, m, to, removed);
        m++;
        c++;
        return false;
      };
      walk(s, 1u, g.cap, leaf);
    } else {
      const uint32_t t = man_tree(g, s);
      c = popc(t);
      if (i < 4u) trees |= t << (8u * i);
    }
    if (i < 4u) packed |= (c < 255u ? c : 255u) << (8u * i);
    i++;
    n += c;
  }
  return n;
}

// k-th move (mover-frame order) of a capture ply with multi-jump origins.
CK_HD Choice select_mixed(const Lane &L, const Gen &g, uint32_t k, uint32_t packed,
                          uint32_t trees, const LeafCache &lc) {
  uint32_t sm[4] = {g.slot[0] & ~g.multi, g.slot[1] & ~g.multi, g.slot[2] & ~g.multi,
                    g.slot[3] & ~g.multi};
  uint32_t o = g.multi, acc = 0, kacc = 0, i = 0;
  while (o) {
    const uint32_t s = ctz(o);
    o &= o - 1u;
    const uint32_t below = bit(s) - 1u;
    const uint32_t start = acc + popc(sm[0] & below) + popc(sm[1] & below) +
                           popc(sm[2] & below) + popc(sm[3] & below);
    if (k < start) break;
    const uint32_t king = getbit(L.kings, s);
    const uint32_t t = king ? 0u : (i < 4u ? (trees >> (8u * i)) & 255u : man_tree(g, s));
    const uint32_t c = i < 4u ? (packed >> (8u * i)) & 255u
                              : (king ? count_walk(L, g, s) : popc(t));
    if (k < start + c) {
      if (!king) return man_leaf(g, s, t, k - start);
      const uint32_t ci = kacc + (k - start);  // index among all king leaves
      if (ci < 4u) {
        Choice ch;
        ch.from = s;
        cache_get(lc, ci, ch.to, ch.removed);
        return ch;
      }
      return select_in_walk(L, g, s, k - start);
    }
    acc += c;
    kacc += king ? c : 0u;
    i++;
  }
  uint32_t sq, sl;
  select_slot(sm, k - acc, sq, sl);
  return slot_choice(L, g, sq, sl);
}

// Weight comparison a/b vs c/d for non-negative integers with a, c >= 1 (b == 0 is +inf).
CK_HD int cmp_weight(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  uint32_t l = a * d, r = c * b;
  return (l > r) - (l < r);
}

// Outcome of a ply: n moves; the reference index chosen (or 0xFFFF at game end).
struct PlyResult {
  uint32_t n;
  uint32_t chosen;
  uint32_t over;    // 1 if the game ended before this ply
  int32_t winner;   // valid when over
};

// Plays one ply of the random policy with word r (src/playout.cpp:24-27,
// src/player.cpp:13-16), given the position's analysis g.  With TRACE, also returns the list
// digest.
template <bool TRACE>
CK_HD PlyResult ply_random_given(Lane &L, const Gen &g, uint32_t r, uint32_t *list_dig) {
  PlyResult res;
  uint32_t packed = 0, trees = 0;
  LeafCache lc = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
  uint32_t n = g.multi ? count_mixed(L, g, packed, trees, lc) : slot_count(g.slot);
  res.n = n;
  if (TRACE) *list_dig = list_digest(L, g, n);
  if (n == 0u || L.counter >= 50u) {  // src/state.cpp:16-23
    res.over = 1u;
    res.chosen = 0xFFFFu;
    res.winner = L.counter < 50u ? (int32_t)(L.turn ^ 1u) : -1;
    return res;
  }
  uint32_t k = mod_small(r, n);
  res.chosen = k;
  res.over = 0u;
  res.winner = 0;
  uint32_t km = L.turn ? n - 1u - k : k;
  Choice c;
  if (g.multi) {
    c = select_mixed(L, g, km, packed, trees, lc);
  } else {
    uint32_t sq, sl;
    select_slot(g.slot, km, sq, sl);
    c = slot_choice(L, g, sq, sl);
  }
  apply(L, c.from, c.to, c.removed);
  return res;
}

// The deferred random ply (a capture ply with a multi-jump origin) from capture trees in
// closed form, with word r and the fast ply's analysis g.  Sets `deep` and leaves the position
// untouched when the trees cannot decide it (ply_random_given then plays it by the walk).
template <bool TRACE>
CK_HD PlyResult ply_random_trees(Lane &L, const Gen &g, uint32_t r, uint32_t *list_dig,
                                 bool &deep) {
  PlyResult res;
  uint32_t m01, m23, dp;
  const uint32_t n = tree_count(L, g, m01, m23, dp);
  deep = dp != 0u;
  res.n = n;
  res.chosen = 0u;
  res.over = 0u;
  res.winner = 0;
  if (deep) return res;
  if (TRACE) *list_dig = list_digest(L, g, n);
  if (L.counter >= 50u) {  // src/state.cpp:16-23 (a capture ply always has moves)
    res.over = 1u;
    res.chosen = 0xFFFFu;
    res.winner = -1;
    return res;
  }
  const uint32_t k = mod_small(r, n);
  res.chosen = k;
  const Choice c = tree_select(L, g, L.turn ? n - 1u - k : k, m01, m23);
  apply(L, c.from, c.to, c.removed);
  return res;
}

// Plays one ply of the random policy with word r.
template <bool TRACE>
CK_HD PlyResult ply_random(Lane &L, uint32_t r, uint32_t *list_dig) {
  Gen g;
  analyse(L, g);
  return ply_random_given<TRACE>(L, g, r, list_dig);
}

// The common part of a random ply: plays it unless it is a capture ply with a multi-jump origin,
// in which case the position is left untouched and `deferred` is returned set (the caller plays
// it later with ply_random_given, the same word and the analysis left in g).
template <bool TRACE>
CK_HD PlyResult ply_random_fast(Lane &L, uint32_t r, uint32_t *list_dig, bool &deferred, Gen &g) {
  analyse(L, g);
  PlyResult res;
  deferred = g.multi != 0u;
  if (deferred) {
    res.n = 0u;
    res.chosen = 0u;
    res.over = 0u;
    res.winner = 0;
    return res;
  }
  const uint32_t n = slot_count(g.slot);
  res.n = n;
  if (TRACE) *list_dig = list_digest(L, g, n);
  if (n == 0u || L.counter >= 50u) {
    res.over = 1u;
    res.chosen = 0xFFFFu;
    res.winner = L.counter < 50u ? (int32_t)(L.turn ^ 1u) : -1;
    return res;
  }
  const uint32_t k = mod_small(r, n);
  res.chosen = k;
  res.over = 0u;
  res.winner = 0;
  uint32_t sq, sl;
  select_slot(g.slot, L.turn ? n - 1u - k : k, sq, sl);
  const Choice c = slot_choice(L, g, sq, sl);
  apply(L, c.from, c.to, c.removed);
  return res;
}

// The heuristic ply after analyse.  WALK = false promises that no capture sequence is longer
// than one jump (the caller deferred such plies), which drops the tree walks from the code.
template <bool TRACE, bool WALK>
CK_HD PlyResult heuristic_body(Lane &L, const Gen &g, uint32_t r, uint32_t *list_dig) {
  PlyResult res;
  const uint32_t sm = material(L.own, L.kings);
  const uint32_t so = material(L.opp, L.kings);
  uint32_t n = 0, ties = 0, bn = 0, bd = 1;
  uint32_t tslot[4];
  const bool walked = WALK && g.capture && g.multi;  // sequences longer than one jump exist
  if (g.capture && !walked) {
    // Single jumps only: a move's weight depends only on whether it removes a king and whether
    // a man lands on the far row, so there are four classes (sm + 3p) / (so - (k ? 4 : 1)),
    // each a set of slot bits.  The tied best classes give the slots to choose from.
    n = slot_count(g.slot);
    const uint32_t men = L.own & ~L.kings, kng = L.own & L.kings;
    const uint32_t okings = L.opp & L.kings;
    // origins whose jump in direction d removes a king (the piece one step away in d)
    const uint32_t kd0 = g.cap[0] & step3(okings), kd1 = g.cap[1] & step2(okings);
    const uint32_t kd2 = g.cap[2] & step1(okings), kd3 = g.cap[3] & step0(okings);
    // per slot (a man's slot 0 jumps in direction 1, its slot 1 in direction 0)
    const uint32_t kc[4] = {g.slot[0] & ((men & kd1) | (kng & kd0)),
                            g.slot[1] & ((men & kd0) | (kng & kd1)), g.slot[2] & kd2,
                            g.slot[3] & kd3};
    const uint32_t pr[4] = {g.slot[0] & men & ROW5, g.slot[1] & men & ROW5, 0u, 0u};
    uint32_t any[4] = {0u, 0u, 0u, 0u};  // class c = 2 * promotion + king removed
#pragma unroll
    for (int t = 0; t < 4; t++) {
      any[0] |= g.slot[t] & ~kc[t] & ~pr[t];
      any[1] |= kc[t] & ~pr[t];
      any[2] |= pr[t] & ~kc[t];
      any[3] |= kc[t] & pr[t];
    }
    uint32_t cn[4], cd[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      cn[c] = sm + ((c & 2) ? 3u : 0u);
      cd[c] = so - ((c & 1) ? 4u : 1u);  // a class that occurs has that piece to remove
    }
    bool have = false;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (any[c] && (!have || cmp_weight(cn[c], cd[c], bn, bd) > 0)) {
        bn = cn[c];
        bd = cd[c];
        have = true;
      }
    }
    tslot[0] = tslot[1] = tslot[2] = tslot[3] = 0u;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (any[c] && cmp_weight(cn[c], cd[c], bn, bd) == 0) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const uint32_t m = ((c & 1) ? kc[t] : g.slot[t] & ~kc[t]) & ((c & 2) ? pr[t] : ~pr[t]);
          tslot[t] |= m;
        }
      }
    }
    ties = slot_count(tslot);
  } else if (g.capture) {
    uint32_t o = g.origins;
    uint32_t king = 0;
    auto leaf = [&](uint32_t to, uint32_t removed, uint32_t, uint32_t) {
      uint32_t num = sm + ((!king && to >= FAR_ROW_FIRST) ? 3u : 0u);
      uint32_t den = so - material(removed, L.kings);
      int c = cmp_weight(num, den, bn, bd);
      if (c > 0) {
        bn = num;
        bd = den;
        ties = 1u;
      } else if (c == 0) {
        ties++;
      }
      n++;
      return false;
    };
    while (o) {
      uint32_t s = ctz(o);
      o &= o - 1u;
      king = getbit(L.kings, s);
      walk(s, king, g.cap, leaf);
    }
  } else {
    n = slot_count(g.slot);
    const uint32_t men = L.own & ~L.kings;
    uint32_t pa = g.slot[0] & men & ROW6, pb = g.slot[1] & men & ROW6;
    // a promotion weighs (sm + 3) / so > sm / so unless so == 0 (all weights +inf)
    bool promo = (pa | pb) && so != 0u;
    tslot[0] = promo ? pa : g.slot[0];
    tslot[1] = promo ? pb : g.slot[1];
    tslot[2] = promo ? 0u : g.slot[2];
    tslot[3] = promo ? 0u : g.slot[3];
    ties = slot_count(tslot);
  }
  res.n = n;
  if (TRACE) *list_dig = list_digest(L, g, n);
  if (n == 0u || L.counter >= 50u) {
    res.over = 1u;
    res.chosen = 0xFFFFu;
    res.winner = L.counter < 50u ? (int32_t)(L.turn ^ 1u) : -1;
    return res;
  }
  uint32_t k = mod_small(r, ties);
  uint32_t km = L.turn ? ties - 1u - k : k;
  res.over = 0u;
  res.winner = 0;
  Choice c = {0u, 0u, 0u};
  if (walked) {
    uint32_t o = g.origins;
    uint32_t king = 0, left = km, idx = 0;
    bool found = false;
    auto leaf = [&](uint32_t to, uint32_t removed, uint32_t, uint32_t) {
      uint32_t num = sm + ((!king && to >= FAR_ROW_FIRST) ? 3u : 0u);
      uint32_t den = so - material(removed, L.kings);
      if (cmp_weight(num, den, bn, bd) == 0) {
        if (left == 0u) {
          c.to = to;
          c.removed = removed;
          found = true;
          return true;
        }
        left--;
      }
      idx++;
      return false;
    };
    while (o && !found) {
      uint32_t s = ctz(o);
      o &= o - 1u;
      king = getbit(L.kings, s);
      c.from = s;
      walk(s, king, g.cap, leaf);
    }
    // reference index of the chosen move (for the trace record)
    res.chosen = L.turn ? n - 1u - idx : idx;
  } else {
    uint32_t sq, sl;
    select_slot(tslot, km, sq, sl);
    c = slot_choice(L, g, sq, sl);
    if (TRACE) {
      // reference index: rank of (sq, sl) in the full slot order
      uint32_t below = bit(sq) - 1u;
      uint32_t i = popc(g.slot[0] & below) + popc(g.slot[1] & below) + popc(g.slot[2] & below) +
                   popc(g.slot[3] & below);
      for (uint32_t t = 0; t < sl; t++) i += getbit(g.slot[t], sq);
      res.chosen = L.turn ? n - 1u - i : i;
    } else {
      res.chosen = 0u;
    }
  }
  apply(L, c.from, c.to, c.removed);
  return res;
}

// Plays one ply of the heuristic policy with word r (src/heuristicPlayout.cpp:20-42 without
// the noise; ties broken by T[r mod |T|]).
template <bool TRACE>
CK_HD PlyResult ply_heuristic(Lane &L, uint32_t r, uint32_t *list_dig) {
  Gen g;
  analyse(L, g);
  return heuristic_body<TRACE, true>(L, g, r, list_dig);
}

// The heuristic ply given the position's analysis g.
template <bool TRACE>
CK_HD PlyResult ply_heuristic_given(Lane &L, const Gen &g, uint32_t r, uint32_t *list_dig) {
  return heuristic_body<TRACE, true>(L, g, r, list_dig);
}

// The common part of a heuristic ply: plays it unless it is a capture ply with a sequence
// longer than one jump, which is left untouched with `deferred` set (the caller plays it later
// with ply_heuristic_given, the same word and the analysis left in g).
template <bool TRACE>
CK_HD PlyResult ply_heuristic_fast(Lane &L, uint32_t r, uint32_t *list_dig, bool &deferred,
                                   Gen &g) {
  analyse(L, g);
  deferred = g.capture && g.multi;
  if (deferred) {
    PlyResult res;
    res.n = 0u;
    res.chosen = 0u;
    res.over = 0u;
    res.winner = 0;
    return res;
  }
  return heuristic_body<TRACE, false>(L, g, r, list_dig);
}

// ------------------------------------------------------------------------------------------
// noisy heuristic (src/heuristicPlayout.cpp:16,26-36; src/heuristicPlayout.cu:50-62;
// src/heuristic.hpp:5): each move's weight (float)(S_me + d_me) / (S_op + d_op) plus an
// N(0, 0.11) draw; the move of largest sum wins, the first in reference order on equality.
// When every move weighs the same, the draws alone decide, and the argmax of iid draws is
// uniform over the moves: such a ply plays reference index r mod n without drawing (the first
// move when the common weight is +inf, which no draw changes).  The draws are made only when
// the weights differ.
// ------------------------------------------------------------------------------------------
constexpr float NOISE_SIGMA = 0.11f;
constexpr uint32_t NOISE_STEP = 0x9E3779B9u;

CK_HD float f32_from_bits(uint32_t b) { return __builtin_bit_cast(float, b); }
CK_HD uint32_t f32_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

// A pair of standard normal draws from two words by the Box-Muller transform (Box and Muller,
// Ann. Math. Statist. 29 (1958) 610-611):
//   radius = sqrt(-2 ln u), u = ((a >> 8) + 1) 2^-24 in (0, 1]
//   angle  = q pi/2 + x, q = b >> 30, x = (t 2^-24 - 1/2) pi/2, t = (b >> 6) mod 2^24
// ln u = (e - 24) ln 2 + ln m with (a >> 8) + 1 = m 2^e, m in (sqrt(1/2), sqrt(2)], and
// ln m = 2 atanh(s), s = (m - 1)/(m + 1), by its series to s^9; sin and cos of x in
// [-pi/4, pi/4) by their Taylor series to x^9 and x^10.  Only IEEE-754 single additions,
// multiplications, one division and one square root, never contracted, so the host and the
// device round every step alike.
CK_HD void normal_pair(uint32_t a, uint32_t b, float &z0, float &z1) {
#pragma clang fp contract(off)
  const float uf = (float)((a >> 8) + 1u);  // exact: at most 2^24
  const uint32_t ub = f32_bits(uf);
  int e = (int)(ub >> 23) - 127;
  float m = f32_from_bits((ub & 0x007FFFFFu) | 0x3F800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  const float lnm =
      s * (2.0f + s2 * (0.666666667f + s2 * (0.4f + s2 * (0.285714286f + s2 * 0.222222222f))));
  const float lnu = (float)(e - 24) * 0.693147181f + lnm;
  const float rad = __builtin_sqrtf(-2.0f * lnu);
  const float x = ((float)((b >> 6) & 0x00FFFFFFu) * 5.96046448e-8f - 0.5f) * 1.57079633f;
  const float x2 = x * x;
  const float sn =
      x * (1.0f + x2 * (-0.166666667f + x2 * (8.33333333e-3f + x2 * (-1.98412698e-4f +
                                                                     x2 * 2.75573192e-6f))));
  const float cs =
      1.0f + x2 * (-0.5f + x2 * (4.16666667e-2f + x2 * (-1.38888889e-3f +
                                                        x2 * (2.48015873e-5f +
                                                              x2 * -2.75573192e-7f))));
  const uint32_t q = b >> 30;
  const float c = q == 0u ? cs : q == 1u ? -sn : q == 2u ? -cs : sn;
  const float t = q == 0u ? sn : q == 1u ? cs : q == 2u ? -sn : -cs;
  z0 = rad * c;
  z1 = rad * t;
}

// Draws of move m (mover-frame order) in a ply of word r: moves 2j and 2j+1 take the two draws
// of the pair from fmix(r + (2j+1) K) and fmix(r + (2j+2) K), K = 0x9E3779B9, scaled by sigma.
CK_HD void noise_pair(uint32_t r, uint32_t j, float &e0, float &e1) {
#pragma clang fp contract(off)
  float z0, z1;
  normal_pair(fmix(r + (2u * j + 1u) * NOISE_STEP), fmix(r + (2u * j + 2u) * NOISE_STEP), z0, z1);
  e0 = z0 * NOISE_SIGMA;
  e1 = z1 * NOISE_SIGMA;
}

// Running argmax of the noisy weights in mover-frame order.  Player 2's reference order is the
// reverse, so its first maximum in reference order is the last one here (>=).
struct NoisyBest {
  float best;
  float spare;  // second draw of the current pair
  uint32_t m;   // moves seen
  uint32_t pick;
  uint32_t rev;
};

CK_HD bool noisy_take(NoisyBest &nb, uint32_t r, float w) {
#pragma clang fp contract(off)
  float e;
  if ((nb.m & 1u) == 0u) {
    noise_pair(r, nb.m >> 1, e, nb.spare);
  } else {
    e = nb.spare;
  }
  const float v = w + e;
  const bool take = nb.rev ? v >= nb.best : v > nb.best;
  if (take) {
    nb.best = v;
    nb.pick = nb.m;
  }
  nb.m++;
  return take;
}

// Weight classes of a ply without longer capture sequences, as slot masks: kc[t] = the move of
// slot t removes a king, prom = the men whose slot-0/1 move promotes.
CK_HD void weight_classes(const Lane &L, const Gen &g, uint32_t kc[4], uint32_t &prom) {
  const uint32_t men = L.own & ~L.kings;
  kc[0] = kc[1] = kc[2] = kc[3] = 0u;
  if (g.capture) {
    const uint32_t kng = L.own & L.kings;
    const uint32_t okings = L.opp & L.kings;
    const uint32_t kd0 = g.cap[0] & step3(okings), kd1 = g.cap[1] & step2(okings);
    kc[0] = g.slot[0] & ((men & kd1) | (kng & kd0));
    kc[1] = g.slot[1] & ((men & kd0) | (kng & kd1));
    kc[2] = g.slot[2] & g.cap[2] & step1(okings);
    kc[3] = g.slot[3] & g.cap[3] & step0(okings);
  }
  prom = men & (g.capture ? ROW5 : ROW6);
}

// Whether every move of such a ply weighs the same; `inf` when that weight is +inf (the mover
// takes the opponent's last piece, or the opponent has none).
CK_HD bool single_weight(const Lane &L, const Gen &g, uint32_t sm, uint32_t so, bool &inf) {
  uint32_t kc[4], prom;
  weight_classes(L, g, kc, prom);
  const uint32_t n = slot_count(g.slot);
  const uint32_t np = popc(g.slot[0] & prom) + popc(g.slot[1] & prom);
  if (!g.capture) {  // steps: sm / so and, for promotions, (sm + 3) / so
    inf = so == 0u;
    return inf || np == 0u || np == n;
  }
  // single jumps: (sm + 3p) / (so - (k ? 4 : 1)) over the classes that occur
  const uint32_t nk = popc(kc[0]) + popc(kc[1]) + popc(kc[2]) + popc(kc[3]);
  const uint32_t nkp = popc(kc[0] & prom) + popc(kc[1] & prom);
  const uint32_t cnt[4] = {n - nk - np + nkp, nk - nkp, np - nkp, nkp};
  bool have = false, same = true;
  uint32_t bn = 0u, bd = 1u;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    if (!cnt[c]) continue;
    const uint32_t num = sm + ((c & 2) ? 3u : 0u), den = so - ((c & 1) ? 4u : 1u);
    if (!have) {
      bn = num;
      bd = den;
      have = true;
    } else if (cmp_weight(num, den, bn, bd) != 0) {
      same = false;
    }
  }
  inf = bd == 0u;
  return same;
}

// The noisy heuristic ply after analyse (WALK as for heuristic_body).
template <bool TRACE, bool WALK>
CK_HD PlyResult noisy_body(Lane &L, const Gen &g, uint32_t r, uint32_t *list_dig) {
  PlyResult res;
  const uint32_t sm = material(L.own, L.kings);
  const uint32_t so = material(L.opp, L.kings);
  const bool walked = WALK && g.capture && g.multi;
  NoisyBest nb = {-__builtin_inff(), 0.0f, 0u, 0u, L.turn};
  Choice c = {0u, 0u, 0u};
  uint32_t n, km;
  bool single, inf = false;
  if (!walked) {
    n = slot_count(g.slot);
    if (TRACE) *list_dig = list_digest(L, g, n);
    if (n == 0u || L.counter >= 50u) {
      res.n = n;
      res.over = 1u;
      res.chosen = 0xFFFFu;
      res.winner = L.counter < 50u ? (int32_t)(L.turn ^ 1u) : -1;
      return res;
    }
    single = single_weight(L, g, sm, so, inf);
    if (!single) {
      uint32_t kc[4], prom;
      weight_classes(L, g, kc, prom);
      float wc[4];  // by class: bit 0 = a king removed, bit 1 = a man promoted
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint32_t num = sm + ((k & 2) ? 3u : 0u);
        const uint32_t den = so - (g.capture ? ((k & 1) ? 4u : 1u) : 0u);
        wc[k] = (float)num / (float)den;  // only classes that occur are read
      }
      uint32_t any = g.slot[0] | g.slot[1] | g.slot[2] | g.slot[3];
      uint32_t bsq = 0u, bsl = 0u;
      while (any) {
        const uint32_t sq = ctz(any);
        any &= any - 1u;
        for (uint32_t sl = 0; sl < 4u; sl++) {
          if (!getbit(g.slot[sl], sq)) continue;
          const uint32_t k = getbit(kc[sl], sq) | ((sl < 2u ? getbit(prom, sq) : 0u) << 1);
          const float w = k == 0u ? wc[0] : k == 1u ? wc[1] : k == 2u ? wc[2] : wc[3];
          if (noisy_take(nb, r, w)) {
            bsq = sq;
            bsl = sl;
          }
        }
      }
      c = slot_choice(L, g, bsq, bsl);
      km = nb.pick;
    } else {
      const uint32_t k = inf ? 0u : mod_small(r, n);  // reference index
      km = L.turn ? n - 1u - k : k;
      uint32_t sq, sl;
      select_slot(g.slot, km, sq, sl);
      c = slot_choice(L, g, sq, sl);
    }
  } else {
    if (L.counter >= 50u) {
      n = count_walks(L, g);
      if (TRACE) *list_dig = list_digest(L, g, n);
      res.n = n;
      res.over = 1u;
      res.chosen = 0xFFFFu;
      res.winner = -1;
      return res;
    }
    // one walk draws for every sequence and notes whether the weights differ at all
    uint32_t o = g.origins;
    uint32_t king = 0, s = 0, fn = 0, fd = 0;
    single = true;
    auto leaf = [&](uint32_t to, uint32_t removed, uint32_t, uint32_t) {
      const uint32_t num = sm + ((!king && to >= FAR_ROW_FIRST) ? 3u : 0u);
      const uint32_t den = so - material(removed, L.kings);
      if (nb.m == 0u) {
        fn = num;
        fd = den;
      } else if (cmp_weight(num, den, fn, fd) != 0) {
        single = false;
      }
      if (noisy_take(nb, r, (float)num / (float)den)) {
        c.from = s;
        c.to = to;
        c.removed = removed;
      }
      return false;
    };
    while (o) {
      s = ctz(o);
      o &= o - 1u;
      king = getbit(L.kings, s);
      walk(s, king, g.cap, leaf);
    }
    n = nb.m;
    km = nb.pick;
    if (single) {
      const uint32_t k = fd == 0u ? 0u : mod_small(r, n);
      km = L.turn ? n - 1u - k : k;
      c = select_walk(L, g, km);
    }
    if (TRACE) *list_dig = list_digest(L, g, n);
  }
  res.n = n;
  res.over = 0u;
  res.winner = 0;
  res.chosen = TRACE ? (L.turn ? n - 1u - km : km) : 0u;
  apply(L, c.from, c.to, c.removed);
  return res;
}

// Plays one ply of the noisy heuristic with word r.
template <bool TRACE>
CK_HD PlyResult ply_noisy(Lane &L, uint32_t r, uint32_t *list_dig) {
  Gen g;
  analyse(L, g);
  return noisy_body<TRACE, true>(L, g, r, list_dig);
}

// The noisy heuristic ply given the position's analysis g.
template <bool TRACE>
CK_HD PlyResult ply_noisy_given(Lane &L, const Gen &g, uint32_t r, uint32_t *list_dig) {
  return noisy_body<TRACE, true>(L, g, r, list_dig);
}

// The common part of a noisy heuristic ply: plays it when its moves all weigh the same (and no
// capture sequence is longer than one jump), else leaves it untouched with `deferred` set, for
// ply_noisy_given with the same word and the analysis left in g.
template <bool TRACE>
CK_HD PlyResult ply_noisy_fast(Lane &L, uint32_t r, uint32_t *list_dig, bool &deferred, Gen &g) {
  PlyResult res;
  analyse(L, g);
  const uint32_t n = slot_count(g.slot);
  bool inf = false;
  deferred = (g.capture && g.multi) ||
             (n != 0u && L.counter < 50u &&
              !single_weight(L, g, material(L.own, L.kings), material(L.opp, L.kings), inf));
  if (deferred) {
    res.n = 0u;
    res.chosen = 0u;
    res.over = 0u;
    res.winner = 0;
    return res;
  }
  res.n = n;
  if (TRACE) *list_dig = list_digest(L, g, n);
  if (n == 0u || L.counter >= 50u) {
    res.over = 1u;
    res.chosen = 0xFFFFu;
    res.winner = L.counter < 50u ? (int32_t)(L.turn ^ 1u) : -1;
    return res;
  }
  const uint32_t k = inf ? 0u : mod_small(r, n);  // reference index
  uint32_t sq, sl;
  select_slot(g.slot, L.turn ? n - 1u - k : k, sq, sl);
  const Choice c = slot_choice(L, g, sq, sl);
  res.over = 0u;
  res.winner = 0;
  res.chosen = k;
  apply(L, c.from, c.to, c.removed);
  return res;
}

// True when the position is finished (no legal move, or the draw counter reached 50).
CK_HD bool game_over(const Lane &L) {
  if (L.counter >= 50u) return true;
  Gen g;
  analyse(L, g);
  return g.capture ? false : slot_count(g.slot) == 0u;
}

}  // namespace ck
