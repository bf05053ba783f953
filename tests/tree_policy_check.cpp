// tree_policy_check.cpp -- how the batch UCB1 tree (gpu_ai_amd/host/mcts.cpp) hands out trials
// when one root move is rigged to win.
//
// From the starting position (seven moves for player 1), every trial below the first root child
// is scored a win for player 1, the mover at the root; every other trial is a fair coin from a
// fixed generator.  After the search, the program prints the trials each root child received
// and which child the tree would play, for the reference's exploitation term (the win fraction
// of the side to move in the child, src/mcts.cpp:189) and for the mover's own win fraction:
//
//   tree_policy_check ITERATIONS TRIALS_PER_ITERATION K PARALLEL
//
// Output lines: "<policy> child <i> trials <n>" and "<policy> opt <i>".
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mcts.hpp"
#include "state.hpp"

namespace {

uint32_t next_bit(uint64_t &s) {
  s = s * 6364136223846793005ull + 1442695040888963407ull;
  return (uint32_t)(s >> 63);
}

void run(const char *name, bool moverWins, unsigned iterations, unsigned trials, double k,
         bool parallel) {
  GameTree tree(getStartingState());
  tree.setExploration(k);
  tree.setMoverWins(moverWins);
  tree.setParallel(parallel);
  uint64_t rng = 12345;
  for (unsigned it = 0; it < iterations; it++) {
    const std::vector<State> leaves = tree.select(trials);
    if (leaves.size() != trials) {
      std::printf("%s: select returned %zu of %u trials\n", name, leaves.size(), trials);
      std::exit(1);
    }
    // the first root child's trials come first (trials are listed child by child)
    const unsigned rigged = tree.isExpanded() ? tree.getChildren()[0]->getAssignedTrials() : 0;
    std::vector<PlayerId> results(leaves.size());
    for (size_t i = 0; i < leaves.size(); i++)
      results[i] = i < rigged ? PLAYER_1 : (next_bit(rng) ? PLAYER_1 : PLAYER_2);
    tree.update(results);
  }
  const auto &children = tree.getChildren();
  const std::vector<Move> &moves = tree.getMoves();
  for (size_t i = 0; i < children.size(); i++)
    std::printf("%s child %zu trials %u\n", name, i, children[i]->getTotalTrials());
  const Move opt = tree.getOptMove(PLAYER_1);
  for (size_t i = 0; i < moves.size(); i++)
    if (moves[i] == opt) std::printf("%s opt %zu\n", name, i);
}

}  // namespace

int main(int argc, char **argv) {
  const unsigned iterations = argc > 1 ? (unsigned)std::atoi(argv[1]) : 200;
  const unsigned trials = argc > 2 ? (unsigned)std::atoi(argv[2]) : 64;
  const double k = argc > 3 ? std::atof(argv[3]) : 2.0;
  const bool parallel = argc > 4 && std::atoi(argv[4]) != 0;
  run("reference", false, iterations, trials, k, parallel);
  run("mover", true, iterations, trials, k, parallel);
  return 0;
}
