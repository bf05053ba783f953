// lane_logic_check.cpp -- the playout kernels' per-lane game logic (gpu_ai_amd/csrc/checkers.h),
// compiled for the host and checked against the oracle (test infrastructure).
//
// Each playout runs the sequence k_playouts runs for one lane: the ply's Philox word, the fast
// ply (ply_random_fast / ply_heuristic_fast / ply_noisy_fast), and, when that defers a
// multi-jump capture, the full ply (ply_random_given / ply_heuristic_given / ply_noisy_given)
// with the same word and the fast ply's
// analysis of the position.  Winners, ply counts, the largest
// list length and the per-ply hash chain over every move list and chosen index must equal the
// oracle's traces (oracle/oracle.c, which follows src/state.cu and src/playout.cpp:17-32).
// Move lists of positions met along the way are compared entry by entry with or_gen_moves.
//
// Usage: lane_logic_check [positions_per_set] [playouts_per_position]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "checkers.h"
#include "oracle.h"

using namespace ck;

namespace {

int play(const or_pos &p, uint32_t seed, uint32_t pi, uint32_t j, uint32_t epoch, int policy,
         or_trace &tr, long &deferred) {
  Lane L = lane_from_abs(p.p1, p.p2, p.kings, p.turn, p.counter);
  uint32_t h = 0, maxn = 0, w[4] = {0, 0, 0, 0};
  for (uint32_t ply = 0;; ply++) {
    if ((ply & 3u) == 0u) philox(ply >> 2, j, TAG_PLAYOUT, epoch, seed, pi, w);
    const uint32_t r = w[ply & 3u];
    uint32_t ld = 0;
    bool d = false;
    Gen g;
    PlyResult res = policy == OR_POLICY_RANDOM      ? ply_random_fast<true>(L, r, &ld, d, g)
                    : policy == OR_POLICY_HEURISTIC ? ply_heuristic_fast<true>(L, r, &ld, d, g)
                                                    : ply_noisy_fast<true>(L, r, &ld, d, g);
    if (d) {
      deferred++;
      // the kernel's walk phase: the full ply with the same word and the fast ply's analysis
      // (random policy: capture trees in closed form, the walk for deep king trees)
      if (policy == OR_POLICY_RANDOM) {
        bool deep = false;
        res = ply_random_trees<true>(L, g, r, &ld, deep);
        if (deep) res = ply_random_given<true>(L, g, r, &ld);
      } else if (policy == OR_POLICY_HEURISTIC) {
        res = ply_heuristic_given<true>(L, g, r, &ld);
      } else {
        res = ply_noisy_given<true>(L, g, r, &ld);
      }
    }
    h = hash_step(h, ply_record(ld, res.n, res.chosen));
    if (res.n > maxn) maxn = res.n;
    if (res.over) {
      tr.winner = (int8_t)res.winner;
      tr.plies = ply;
      tr.hash = h;
      tr.max_n = (uint8_t)(maxn > 255u ? 255u : maxn);
      return res.winner;
    }
  }
}

// Move list of p through the lane enumerator, in the reference's order and absolute squares.
int lane_moves(const or_pos &p, or_move *out, int cap) {
  const Lane L = lane_from_abs(p.p1, p.p2, p.kings, p.turn, p.counter);
  Gen g;
  analyse(L, g);
  const uint32_t n = g.capture ? count_walks(L, g) : slot_count(g.slot);
  if ((int)n > cap) return -1;
  int bad = 0;
  auto visit = [&](uint32_t idx, uint32_t from, uint32_t to, uint32_t removed, uint32_t jumps,
                   uint32_t dirs, uint32_t king) {
    (void)removed;
    const uint32_t ri = L.turn ? n - 1u - idx : idx;
    if (ri >= n) {
      bad = 1;
      return;
    }
    or_move m;
    std::memset(&m, 0, sizeof m);
    m.from = (uint8_t)abs_sq(L, from);
    m.to = (uint8_t)abs_sq(L, to);
    m.jumps = (uint8_t)jumps;
    m.promoted = (uint8_t)(!king && to >= FAR_ROW_FIRST);
    uint32_t x = from;
    for (uint32_t t = 0; t < jumps; t++) {
      const uint32_t d = (dirs >> (2u * t)) & 3u;
      const uint32_t land = (uint32_t)((int)x + jump_off(d));
      m.inter[t] = (uint8_t)abs_sq(L, land);
      m.removed[t] = (uint8_t)abs_sq(L, mid_sq(x, land));
      x = land;
    }
    out[ri] = m;
  };
  enumerate(L, g, visit);
  return bad ? -1 : (int)n;
}

// The deferred ply's move choice (count_mixed / select_mixed: men's trees in closed form,
// kings' by the walk) against the walk's enumeration, for every move index of a capture
// position with a multi-jump origin.  Returns the number of disagreeing indices.
long deep_trees = 0;  // multi-jump positions whose trees the closed forms leave to the walk

long check_mixed(const or_pos &p, long &checked) {
  const Lane L = lane_from_abs(p.p1, p.p2, p.kings, p.turn, p.counter);
  Gen g;
  analyse(L, g);
  if (!g.capture || !g.multi) return 0;
  uint32_t packed = 0, trees = 0;
  LeafCache lc = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
  const uint32_t n = count_mixed(L, g, packed, trees, lc);
  long bad = n != count_walks(L, g);
  uint32_t from[OR_MAX_MOVES], to[OR_MAX_MOVES], rm[OR_MAX_MOVES];
  uint32_t m = 0;
  auto visit = [&](uint32_t idx, uint32_t f, uint32_t t, uint32_t r, uint32_t, uint32_t,
                   uint32_t) {
    if (idx < OR_MAX_MOVES) {
      from[idx] = f;
      to[idx] = t;
      rm[idx] = r;
    }
    m++;
  };
  enumerate(L, g, visit);
  for (uint32_t k = 0; k < n && k < m && k < OR_MAX_MOVES; k++) {
    const Choice c = select_mixed(L, g, k, packed, trees, lc);
    checked++;
    bad += c.from != from[k] || c.to != to[k] || c.removed != rm[k];
  }
  // the same through the capture trees of every origin, where they decide the ply
  uint32_t m01, m23, deep;
  const uint32_t nt = tree_count(L, g, m01, m23, deep);
  // the closed forms give up exactly when a king has a sequence of three or more jumps (or more
  // than four origins can capture)
  uint32_t maxd = 0;
  for (uint32_t o = g.origins & L.kings; o; o &= o - 1u) {
    auto leaf = [&](uint32_t, uint32_t, uint32_t d, uint32_t) {
      maxd = d > maxd ? d : maxd;
      return false;
    };
    walk(ctz(o), 1u, g.cap, leaf);
  }
  bad += (deep != 0u) != (maxd >= 3u || popc(g.origins) > 4u);
  if (!deep) {
    bad += nt != n;
    for (uint32_t k = 0; k < nt && k < m && k < OR_MAX_MOVES; k++) {
      const Choice c = tree_select(L, g, k, m01, m23);
      checked++;
      bad += c.from != from[k] || c.to != to[k] || c.removed != rm[k];
    }
  } else {
    deep_trees++;
  }
  return bad;
}

// The k-th move of the slot masks by a plain binary search over squares (four masked popcounts
// per step), against which select_slot's byte-transposed search is checked.
void select_slot_plain(const uint32_t s[4], uint32_t k, uint32_t &sq, uint32_t &sl) {
  uint32_t lo = 0;
  for (uint32_t stride = 16; stride >= 1; stride >>= 1) {
    const uint32_t below = bit(lo + stride) - 1u;
    if (popc(s[0] & below) + popc(s[1] & below) + popc(s[2] & below) + popc(s[3] & below) <= k)
      lo += stride;
  }
  const uint32_t below = bit(lo) - 1u;
  uint32_t j = k - (popc(s[0] & below) + popc(s[1] & below) + popc(s[2] & below) +
                    popc(s[3] & below));
  for (sl = 0; sl < 4u; sl++)
    if (getbit(s[sl], lo) && j-- == 0u) break;
  sq = lo;
}

// select_slot against the plain search for every k of random slot masks.
long check_select_slot(int masks, long &checked) {
  long bad = 0;
  uint64_t x = 12345;
  auto next = [&]() {
    x = x * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(x >> 32);
  };
  for (int i = 0; i < masks; i++) {
    const uint32_t dens = next() % 4u;
    uint32_t s[4];
    for (int t = 0; t < 4; t++) {
      s[t] = next();
      for (uint32_t d = 0; d < dens; d++) s[t] &= next();
    }
    const uint32_t n = popc(s[0]) + popc(s[1]) + popc(s[2]) + popc(s[3]);
    for (uint32_t k = 0; k < n; k++) {
      uint32_t a, b, c, d;
      select_slot(s, k, a, b);
      select_slot_plain(s, k, c, d);
      bad += a != c || b != d;
      checked++;
    }
  }
  return bad;
}

// Random placements of men and kings (dense enough for long capture trees), from a fixed LCG.
std::vector<or_pos> placements(int n) {
  std::vector<or_pos> v((size_t)n);
  uint64_t x = 0x9E3779B97F4A7C15ull;
  auto next = [&]() {
    x = x * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(x >> 32);
  };
  for (int i = 0; i < n; i++) {
    const uint32_t occ = next() & next() | (next() & next() & next());  // about 44% occupied
    const uint32_t side = next();
    const uint32_t king = next() & next() & next();  // about 12% kings
    v[i].p1 = occ & side;
    v[i].p2 = occ & ~side;
    v[i].kings = occ & king;
    v[i].turn = (uint8_t)(next() & 1u);
    v[i].counter = 0;
    v[i].pad[0] = v[i].pad[1] = 0;
  }
  return v;
}

}  // namespace

int main(int argc, char **argv) {
  const int N = argc > 1 ? std::atoi(argv[1]) : 1024;
  const int P = argc > 2 ? std::atoi(argv[2]) : 8;
  long bad = 0, playouts = 0, deferred = 0, lists = 0, mixed = 0;
  for (int set = 0; set < 4; set++) {
    std::vector<or_pos> pos(N);
    if (set == 0) or_gen_start_positions(2016, N, 0, 30, 0, pos.data());
    else if (set == 1) or_gen_start_positions(777, N, 20, 60, 1, pos.data());
    else if (set == 2) or_gen_start_positions(99, N, 40, 120, 1, pos.data());
    else {  // random placements, draw counters near the limit: finished games, captures of the
            // last piece (heuristic weight +inf), long king sequences
      pos = placements(N);
      for (int i = 0; i < N; i++) pos[i].counter = (uint8_t)(i % 4 == 0 ? 48 : i % 4 == 1 ? 49 : 0);
    }
    for (int policy = 0; policy < 3; policy++) {
      std::vector<or_trace> ot((size_t)N * P);
      or_run_playouts(pos.data(), N, nullptr, P, 12345, 3, policy, nullptr, ot.data(), 0);
      long pbad = 0;
      for (int i = 0; i < N; i++)
        for (int j = 0; j < P; j++) {
          or_trace t;
          play(pos[i], 12345, i, j, 3, policy, t, deferred);
          const or_trace &o = ot[(size_t)i * P + j];
          playouts++;
          if (t.winner != o.winner || t.plies != o.plies || t.hash != o.hash || t.max_n != o.max_n) {
            if (pbad < 5)
              std::printf("playout mismatch: set %d policy %d position %d playout %d\n", set,
                          policy, i, j);
            pbad++;
          }
        }
      bad += pbad;
    }
    // move lists along a deterministic walk from each position
    for (int i = 0; i < N; i++) {
      or_pos p = pos[i];
      for (int ply = 0; ply < 300; ply++) {
        or_move om[OR_MAX_MOVES], lm[OR_MAX_MOVES];
        const int n = or_gen_moves(&p, om, OR_MAX_MOVES);
        const int m = lane_moves(p, lm, OR_MAX_MOVES);
        lists++;
        if (n != m || (n > 0 && std::memcmp(om, lm, sizeof(or_move) * (size_t)n) != 0)) {
          if (bad < 5) std::printf("move list mismatch: set %d position %d ply %d\n", set, i, ply);
          bad++;
        }
        const long mb = check_mixed(p, mixed);
        if (mb && bad < 5) std::printf("deferred-ply choice mismatch: set %d position %d ply %d\n",
                                       set, i, ply);
        bad += mb;
        if (n <= 0 || p.counter >= OR_DRAW_MOVES) break;
        or_apply(&p, &om[(ply * 7 + i * 3) % n]);
      }
    }
  }
  for (const or_pos &p : placements(400 * N)) {
    const long mb = check_mixed(p, mixed);
    if (mb && bad < 5)
      std::printf("deferred-ply choice mismatch: placement %08x %08x %08x turn %d\n", p.p1, p.p2,
                  p.kings, p.turn);
    bad += mb;
  }
  long slots = 0;
  const long sb = check_select_slot(20000, slots);
  if (sb) std::printf("select_slot disagrees with the plain search on %ld of %ld choices\n", sb,
                      slots);
  bad += sb;
  std::printf("playouts %ld (deferred plies %ld), move lists %ld, deferred choices %ld "
              "(deep trees %ld), mismatches %ld\n", playouts, deferred, lists, mixed, deep_trees,
              bad);
  return bad != 0;
}
