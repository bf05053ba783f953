// sanitize_check.cpp -- host-only runs of the search code for the address, undefined-behaviour
// and thread sanitizers (tests/test_sanitizers.py builds it three ways).
//
//   sanitize_check player   MCTSPlayer (gpu_ai_amd/host/player.cpp): its worker thread searching
//                           the batch UCB1 tree (OpenMP selection at the root) while the game
//                           thread calls getMove and move under the tree mutex, for a few plies
//                           of self-play, then stop
//   sanitize_check tree     GameTree select/update with OpenMP at the root, several trees
//                           searched from concurrent threads (as the multi-search does)
//
// Leaves are scored by HashDriver, a PlayoutDriver that needs no device: the winner of a leaf is
// a hash of its position and of a per-call counter, so the searches take real, varied paths.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "mcts.hpp"
#include "player.hpp"
#include "playout.hpp"
#include "state.hpp"

namespace {

uint32_t mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}

class HashDriver : public PlayoutDriver {
 public:
  std::vector<PlayerId> runPlayouts(std::vector<State> states) override {
    const uint32_t call = calls_.fetch_add(1u);
    std::vector<PlayerId> out(states.size());
    for (size_t i = 0; i < states.size(); i++) {
      const mi_pos p = states[i].pack();
      const uint32_t h = mix(p.p1 ^ mix(p.p2 ^ mix(p.kings ^ p.turn)) ^ mix(call * 0x9E3779B9u + (uint32_t)i));
      out[i] = states[i].isGameOver() ? states[i].getWinner() : (PlayerId)((int)(h % 3u) - 1);
    }
    return out;
  }
  std::string getName() const override { return "hash"; }

 private:
  std::atomic<uint32_t> calls_{0};
};

int runPlayer() {
  MCTSPlayer a(64, 0.25f, 0.02, std::make_unique<HashDriver>());
  MCTSPlayer b(64, 0.25f, 0.02, std::make_unique<HashDriver>());
  State s = getStartingState();
  a.start();
  b.start();
  for (int ply = 0; ply < 12 && !s.isGameOver(); ply++) {
    MCTSPlayer &mover = s.turn == PLAYER_1 ? a : b;
    const Move m = mover.getMove(s, false);
    a.move(m);
    b.move(m);
    s.move(m);
  }
  a.stop();
  b.stop();
  const SearchStats sa = a.stats(), sb = b.stats();
  std::printf("player: iterations %llu + %llu, playouts %llu + %llu\n",
              (unsigned long long)sa.iterations, (unsigned long long)sb.iterations,
              (unsigned long long)sa.playouts, (unsigned long long)sb.playouts);
  return sa.iterations > 0 && sb.iterations > 0 ? 0 : 1;
}

int runTrees() {
  const int threads = 4;
  std::vector<std::thread> pool;
  std::vector<unsigned> trials(threads);
  for (int k = 0; k < threads; k++) {
    pool.emplace_back([k, &trials] {
      HashDriver d;
      GameTree tree(getStartingState());
      tree.setParallel(k == 0);  // one tree selects with OpenMP at its root
      tree.setExploration(k % 2 ? 2.0 : 16.0);
      for (int it = 0; it < 60; it++) tree.expand(96, &d);
      trials[k] = tree.getTotalTrials();
    });
  }
  for (auto &t : pool) t.join();
  for (int k = 0; k < threads; k++) std::printf("tree %d: %u trials\n", k, trials[k]);
  for (unsigned t : trials)
    if (t != 60u * 96u) return 1;
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "player") == 0) return runPlayer();
  if (argc > 1 && std::strcmp(argv[1], "tree") == 0) return runTrees();
  std::fprintf(stderr, "usage: sanitize_check player|tree\n");
  return 2;
}
