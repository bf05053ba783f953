// ref_harness.cpp -- C entry points over the reference's own host code (test infrastructure).
//
// Built only in the development container by oracle/Makefile, against the unmodified
// reference sources src/state.cu, src/state.cpp, src/heuristic.cu, src/heuristicPlayout.cpp and
// src/mcts.cpp, which the Makefile compiles where they lie.  The functions below drive that
// code with the same packed positions, Philox streams and hash records as oracle.c, so that the
// restatement can be compared with the reference on any input.  The driver loops restate the
// ten-line bodies of HostPlayoutDriver::runPlayouts (src/playout.cpp:17-32) and
// HostHeuristicPlayoutDriver::runPlayouts (src/heuristicPlayout.cpp:12-48), replacing
// rand() % n and the N(0, 0.11) noise by the Philox word of the ply (the noisy policy draws its
// noise with or_heuristic_noise, and plays r mod n where all weights are equal, where the argmax
// of iid draws is uniform); ref_host_heuristic_driver runs the reference's own
// HostHeuristicPlayoutDriver, noise and all.
#include "state.hpp"
#include "heuristic.hpp"
#include "mcts.hpp"
#include "playout.hpp"

#include "oracle.h"

#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include <omp.h>

namespace {

Loc locOf(int s) {
  uint8_t row = (uint8_t)(s >> 2);
  return Loc(row, (uint8_t)(2 * (s & 3) + ((row & 1) == 0 ? 1 : 0)));
}
int sqOf(Loc l) { return 4 * l.row + (l.col >> 1); }

State toState(const or_pos &p) {
  State st = getStartingState();
  for (int s = 0; s < 32; s++) {
    Loc l = locOf(s);
    BoardItem item = {false, CHECKER, PLAYER_1};
    uint32_t bit = 1u << s;
    if ((p.p1 | p.p2) & bit) {
      item.occupied = true;
      item.owner = (p.p1 & bit) ? PLAYER_1 : PLAYER_2;
      item.type = (p.kings & bit) ? CHECKER_KING : CHECKER;
    }
    st.board[l.row][l.col] = item;
  }
  st.turn = p.turn ? PLAYER_2 : PLAYER_1;
  st.movesSinceLastCapture = p.counter;
  return st;
}

or_pos fromState(const State &st) {
  or_pos p;
  std::memset(&p, 0, sizeof(p));
  for (int s = 0; s < 32; s++) {
    BoardItem item = st[locOf(s)];
    if (!item.occupied) continue;
    if (item.owner == PLAYER_1) p.p1 |= 1u << s; else p.p2 |= 1u << s;
    if (item.type == CHECKER_KING) p.kings |= 1u << s;
  }
  p.turn = st.turn == PLAYER_2 ? 1 : 0;
  p.counter = (uint8_t)(st.movesSinceLastCapture > 255 ? 255 : st.movesSinceLastCapture);
  return p;
}

or_move exportMove(const Move &m) {
  or_move o;
  std::memset(&o, 0, sizeof(o));
  o.from = (uint8_t)sqOf(m.from);
  o.to = (uint8_t)sqOf(m.to);
  o.jumps = m.jumps;
  o.promoted = m.promoted ? 1 : 0;
  for (int i = 0; i < m.jumps; i++) {
    o.inter[i] = (uint8_t)sqOf(m.intermediate[i]);
    o.removed[i] = (uint8_t)sqOf(m.removed[i]);
  }
  return o;
}

uint32_t listDigest(const std::vector<Move> &moves) {
  uint32_t acc = 0;
  for (size_t i = 0; i < moves.size(); i++) {
    or_move o = exportMove(moves[i]);
    acc += or_fmix32(or_move_digest(&o) + (uint32_t)i * 0x9E3779B9u);
  }
  return acc;
}

struct Words {
  uint32_t w[4];
  void refill(uint32_t seed, uint32_t pos, uint32_t c0, uint32_t c1, uint32_t tag, uint32_t c3) {
    uint32_t ctr[4] = {c0, c1, tag, c3}, key[2] = {seed, pos};
    or_philox4x32_10(ctr, key, w);
  }
};

// One playout in the shape of the reference host drivers.
int playout(State state, uint32_t seed, uint32_t posIndex, uint32_t playoutIndex, uint32_t epoch,
            int policy, or_trace *tr) {
  uint32_t h = 0, plies = 0, captures = 0;
  int maxN = 0;
  unsigned stateScore[NUM_PLAYERS];
  if (policy == OR_POLICY_HEURISTIC || policy == OR_POLICY_HEURISTIC_NOISY)
    scoreState(state, stateScore);
  Words words;
  for (uint32_t ply = 0;; ply++) {
    bool over = state.isGameOver();            // first genMoves of the ply
    std::vector<Move> moves = state.getMoves(); // second genMoves, as RandomPlayer::getMove
    int n = (int)moves.size();
    if (n > maxN) maxN = n;
    uint32_t L = tr ? listDigest(moves) : 0;
    if (over) {
      if (tr) h = or_hash_step(h, or_ply_record(L, n, -1));
      break;
    }
    unsigned k;
    if (policy == OR_POLICY_RANDOM_LIBC) {
      k = (unsigned)(rand() % (int)moves.size());  // src/player.cpp:15, global-lock rand()
      if (moves[0].jumps > 0) captures++;
      if (tr) h = or_hash_step(h, or_ply_record(L, n, (int)k));
      state.move(moves[k]);
      plies++;
      continue;
    }
    if ((ply & 3) == 0) words.refill(seed, posIndex, ply >> 2, playoutIndex, OR_TAG_PLAYOUT, epoch);
    uint32_t r = words.w[ply & 3];
    if (policy == OR_POLICY_HEURISTIC) {
      std::vector<std::array<int, NUM_PLAYERS>> moveScores(moves.size());
      std::vector<float> weights(moves.size());
      float optWeight = -1 / 0.0;
      for (size_t i = 0; i < moves.size(); i++) {
        scoreMove(state, moves[i], moveScores[i].data());
        weights[i] = getWeight(state, stateScore, moveScores[i].data());
        if (weights[i] > optWeight) optWeight = weights[i];
      }
      std::vector<unsigned> ties;
      for (size_t i = 0; i < moves.size(); i++)
        if (weights[i] == optWeight) ties.push_back((unsigned)i);
      k = ties[r % ties.size()];
      for (int i = 0; i < NUM_PLAYERS; i++) stateScore[i] += moveScores[k][i];
    } else if (policy == OR_POLICY_HEURISTIC_NOISY) {
      // src/heuristicPlayout.cpp:26-36 with the draw of move i taken from the ply's word (mover
      // order: player 2's list is generated in the reverse of its mover-frame order); a ply
      // whose weights are all equal plays index r mod n (0 for +inf) without drawing
      std::vector<std::array<int, NUM_PLAYERS>> moveScores(moves.size());
      std::vector<float> weights(moves.size());
      bool single = true;
      for (size_t i = 0; i < moves.size(); i++) {
        scoreMove(state, moves[i], moveScores[i].data());
        weights[i] = getWeight(state, stateScore, moveScores[i].data());
        if (weights[i] != weights[0]) single = false;
      }
      if (single) {
        k = std::isinf(weights[0]) ? 0u : (unsigned)(r % moves.size());
      } else {
        float optWeight = -1 / 0.0;
        k = (unsigned)-1;
        for (size_t i = 0; i < moves.size(); i++) {
          const uint32_t m = state.turn == PLAYER_1 ? (uint32_t)i : (uint32_t)(n - 1 - (int)i);
          float weight = weights[i] + or_heuristic_noise(r, m);
          if (weight > optWeight) {
            k = (unsigned)i;
            optWeight = weight;
          }
        }
      }
      for (int i = 0; i < NUM_PLAYERS; i++) stateScore[i] += moveScores[k][i];
    } else {
      k = r % moves.size();
    }
    if (moves[0].jumps > 0) captures++;
    if (tr) h = or_hash_step(h, or_ply_record(L, n, (int)k));
    state.move(moves[k]);
    plies++;
  }
  PlayerId w = state.getWinner();
  if (tr) {
    tr->winner = (int8_t)w;
    tr->max_n = (uint8_t)(maxN > 255 ? 255 : maxN);
    tr->pad = 0;
    tr->plies = plies;
    tr->hash = h;
    tr->captures = captures;
  }
  return (int)w;
}

uint64_t perft(const State &s, int depth) {
  if (depth == 0) return 1;
  Move result[MAX_MOVES];
  uint8_t n = s.genMoves(result);
  if (depth == 1) return n;
  uint64_t total = 0;
  for (uint8_t i = 0; i < n; i++) {
    State t = s;
    t.move(result[i]);
    total += perft(t, depth - 1);
  }
  return total;
}

}  // namespace

extern "C" {

int ref_gen_moves(const or_pos *p, or_move *out, int cap) {
  State st = toState(*p);
  Move result[MAX_MOVES];
  int n = st.genMoves(result);
  if (n > cap) return -1;
  for (int i = 0; i < n; i++) out[i] = exportMove(result[i]);
  return n;
}

// Applies the index-th generated move through State::move.
int ref_apply_index(or_pos *p, int index) {
  State st = toState(*p);
  Move result[MAX_MOVES];
  int n = st.genMoves(result);
  if (index < 0 || index >= n) return -1;
  st.move(result[index]);
  *p = fromState(st);
  return 0;
}

int ref_is_game_over(const or_pos *p) { return toState(*p).isGameOver() ? 1 : 0; }

int ref_winner(const or_pos *p) { return (int)toState(*p).getWinner(); }

uint64_t ref_perft(const or_pos *p, int depth) { return perft(toState(*p), depth); }

void ref_score_state(const or_pos *p, uint32_t score[2]) {
  unsigned s[NUM_PLAYERS];
  scoreState(toState(*p), s);
  score[0] = s[0];
  score[1] = s[1];
}

// Scores of the index-th generated move (scoreMove) and its weight (getWeight).
int ref_score_move_index(const or_pos *p, int index, int32_t score[2], float *weight) {
  State st = toState(*p);
  Move result[MAX_MOVES];
  int n = st.genMoves(result);
  if (index < 0 || index >= n) return -1;
  int ms[NUM_PLAYERS];
  unsigned ss[NUM_PLAYERS];
  scoreState(st, ss);
  scoreMove(st, result[index], ms);
  score[0] = ms[0];
  score[1] = ms[1];
  *weight = getWeight(st, ss, ms);
  return 0;
}

int ref_playout(const or_pos *p, uint32_t seed, uint32_t posIndex, uint32_t playoutIndex,
                uint32_t epoch, int policy, or_trace *tr) {
  return playout(toState(*p), seed, posIndex, playoutIndex, epoch, policy, tr);
}

int ref_run_playouts(const or_pos *pos, uint32_t npos, const uint32_t *counts, uint32_t perPos,
                     uint32_t seed, uint32_t epoch, int policy, int8_t *winners, or_trace *traces,
                     int threads) {
  std::vector<uint64_t> offs(npos + 1, 0);
  for (uint32_t i = 0; i < npos; i++) offs[i + 1] = offs[i] + (counts ? counts[i] : perPos);
  std::vector<State> states(npos);
  for (uint32_t i = 0; i < npos; i++) states[i] = toState(pos[i]);
  int64_t total = (int64_t)offs[npos];
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t g = 0; g < total; g++) {
    uint32_t lo = 0, hi = npos;
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) / 2;
      if (offs[mid] <= (uint64_t)g) lo = mid; else hi = mid;
    }
    or_trace t;
    int w = playout(states[lo], seed, lo, (uint32_t)((uint64_t)g - offs[lo]), epoch, policy,
                    traces ? &t : nullptr);
    if (winners) winners[g] = (int8_t)w;
    if (traces) traces[g] = t;
  }
  return 0;
}

// The reference's own HostHeuristicPlayoutDriver::runPlayouts (src/heuristicPlayout.cpp:12-48)
// on per_pos copies of each position, position-major, on one OpenMP thread: its generator is a
// default-seeded std::default_random_engine shared by the loop's threads, so one thread makes
// the call reproducible, and the copies of a position draw successive noise from that stream.
int ref_host_heuristic_driver(const or_pos *pos, uint32_t npos, uint32_t perPos, int8_t *winners) {
  std::vector<State> states;
  states.reserve((size_t)npos * perPos);
  for (uint32_t i = 0; i < npos; i++) {
    State st = toState(pos[i]);
    for (uint32_t j = 0; j < perPos; j++) states.push_back(st);
  }
  const int saved = omp_get_max_threads();
  omp_set_num_threads(1);
  HostHeuristicPlayoutDriver driver;
  std::vector<PlayerId> w = driver.runPlayouts(states);
  omp_set_num_threads(saved);
  for (size_t i = 0; i < w.size(); i++) winners[i] = (int8_t)w[i];
  return 0;
}

// Choice frequencies of one noisy heuristic ply from the position (src/heuristicPlayout.cpp:
// 15-16,22-36 for one ply, drawn `trials` times): the reference's weights and its own noise
// source, std::normal_distribution<float>(0, HEURISTIC_SIGMA) over a std::default_random_engine
// seeded with `seed`.  counts[i] = times the move of reference index i was chosen; returns n.
int ref_noisy_choice_counts(const or_pos *p, uint32_t trials, uint32_t seed, uint32_t *counts,
                            int cap) {
  State state = toState(*p);
  std::vector<Move> moves = state.getMoves();
  if ((int)moves.size() > cap) return -1;
  unsigned stateScore[NUM_PLAYERS];
  scoreState(state, stateScore);
  std::vector<float> weights(moves.size());
  for (size_t i = 0; i < moves.size(); i++) {
    int moveScore[NUM_PLAYERS];
    scoreMove(state, moves[i], moveScore);
    weights[i] = getWeight(state, stateScore, moveScore);
  }
  std::default_random_engine generator(seed);
  std::normal_distribution<float> distribution(0, HEURISTIC_SIGMA);
  for (size_t i = 0; i < moves.size(); i++) counts[i] = 0;
  for (uint32_t t = 0; t < trials; t++) {
    unsigned optIndex = (unsigned)-1;
    float optWeight = -1 / 0.0;
    for (size_t i = 0; i < moves.size(); i++) {
      float weight = weights[i] + distribution(generator);
      if (weight > optWeight) {
        optIndex = (unsigned)i;
        optWeight = weight;
      }
    }
    counts[optIndex]++;
  }
  return (int)moves.size();
}

// Start positions in the manner of genRandomStates (src/driver.cpp:76-104), with the same
// stream layout as or_gen_start_positions.
int ref_gen_start_positions(uint32_t seed, uint32_t n, uint32_t kmin, uint32_t kmax,
                            int rejectTerminal, or_pos *out) {
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint32_t posIndex = (uint32_t)i;
    State state = getStartingState();
    bool done = false;
    for (uint32_t attempt = 0; attempt < 4096 && !done; attempt++) {
      state = getStartingState();
      Words len;
      len.refill(seed, posIndex, 0, attempt, OR_TAG_START, 1);
      uint32_t k = kmin + len.w[0] % (kmax - kmin + 1);
      Words words;
      for (uint32_t m = 0; m < k; m++) {
        if (state.isGameOver()) break;
        if ((m & 3) == 0) words.refill(seed, posIndex, m >> 2, attempt, OR_TAG_START, 0);
        std::vector<Move> moves = state.getMoves();
        state.move(moves[words.w[m & 3] % moves.size()]);
      }
      if (!rejectTerminal || !state.isGameOver()) done = true;
    }
    if (!done) {
#pragma omp atomic write
      failed = 1;
    }
    out[posIndex] = fromState(state);
  }
  return failed ? -1 : 0;
}

// The trace of gpuai_mcts_trace (include/gpuai/gpuai_host.h) over the reference's own GameTree
// (src/mcts.cpp): identical trial counts and winner function, so the two traces must agree.
int ref_mcts_trace(const or_pos *root, uint32_t iterations, uint32_t initial, double scale,
                   uint32_t *trace, uint32_t traceCap) {
  if (traceCap < 5u * iterations + 4u) return -1;
  auto rotl = [](uint32_t x, int r) { return (x << r) | (x >> (32 - r)); };
  GameTree tree(toState(*root));
  uint32_t w = 0;
  for (uint32_t it = 0; it < iterations; it++) {
    unsigned n = (unsigned)(tree.getTotalTrials() * scale);
    if (n == 0) n = initial;
    std::vector<State> leaves = tree.select(n);
    std::vector<PlayerId> results(leaves.size());
    uint32_t digest = 0;
    for (size_t i = 0; i < leaves.size(); i++) {
      const or_pos p = fromState(leaves[i]);
      const uint32_t key = p.p1 ^ rotl(p.p2, 7) ^ rotl(p.kings, 13) ^ ((uint32_t)p.turn << 30) ^
                           ((uint32_t)p.counter << 20);
      digest = or_fmix32(digest ^ key) + 0x9E3779B9u;
      const uint32_t h = or_fmix32(key ^ (it * 0x9E3779B9u) ^ ((uint32_t)i * 0x85EBCA6Bu));
      results[i] = (PlayerId)((int)(h % 3u) - 1);
    }
    tree.update(results);
    trace[w++] = (uint32_t)leaves.size();
    trace[w++] = digest;
    const unsigned total = tree.getTotalTrials();
    trace[w++] = total;
    // wins are private in the reference; getScore(p) = wins[p] / totalTrials recovers them
    trace[w++] = (uint32_t)std::llround(tree.getScore(PLAYER_1) * total);
    trace[w++] = (uint32_t)std::llround(tree.getScore(PLAYER_2) * total);
  }
  for (PlayerId p : {PLAYER_1, PLAYER_2}) {
    Move m = tree.getOptMove(p);
    trace[w++] = (uint32_t)sqOf(m.from);
    trace[w++] = (uint32_t)sqOf(m.to);
  }
  return (int)w;
}

}  // extern "C"
