// lane_logic_check.cpp -- the playout kernels' per-lane game logic (gpu_ai_amd/csrc/checkers.h),
// compiled for the host and checked against the oracle (test infrastructure).
//
// Each playout runs the sequence k_playouts runs for one lane: the ply's Philox word, the fast
// ply (ply_random_fast / ply_heuristic_fast), and, when that defers a multi-jump capture, the
// full ply (ply_random_given / ply_heuristic_given) with the same word and the fast ply's
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
    PlyResult res = policy == OR_POLICY_RANDOM ? ply_random_fast<true>(L, r, &ld, d, g)
                                               : ply_heuristic_fast<true>(L, r, &ld, d, g);
    if (d) {
      deferred++;
      // the kernel's walk phase: the full ply with the same word and the fast ply's analysis
      res = policy == OR_POLICY_RANDOM ? ply_random_given<true>(L, g, r, &ld)
                                       : ply_heuristic_given<true>(L, g, r, &ld);
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

}  // namespace

int main(int argc, char **argv) {
  const int N = argc > 1 ? std::atoi(argv[1]) : 1024;
  const int P = argc > 2 ? std::atoi(argv[2]) : 8;
  long bad = 0, playouts = 0, deferred = 0, lists = 0;
  for (int set = 0; set < 3; set++) {
    std::vector<or_pos> pos(N);
    if (set == 0) or_gen_start_positions(2016, N, 0, 30, 0, pos.data());
    else if (set == 1) or_gen_start_positions(777, N, 20, 60, 1, pos.data());
    else or_gen_start_positions(99, N, 40, 120, 1, pos.data());
    for (int policy = 0; policy < 2; policy++) {
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
        if (n <= 0 || p.counter >= OR_DRAW_MOVES) break;
        or_apply(&p, &om[(ply * 7 + i * 3) % n]);
      }
    }
  }
  std::printf("playouts %ld (deferred plies %ld), move lists %ld, mismatches %ld\n", playouts,
              deferred, lists, bad);
  return bad != 0;
}
