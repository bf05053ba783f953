// netmatch.hpp -- value-network leaves against playout leaves: a match between two MCTS players
// that differ only in how a leaf is scored.
//
// Both players are the reference's batch UCB1 search (GameTree, src/mcts.cpp) keeping their tree
// across moves as MCTSPlayer does (src/player.cpp:111-116).  The playout player scores each leaf
// by one random playout on the device (src/mcts.cpp:57-61); the network player by one batched
// value-network evaluation (finished games by their result).  Games are played in lockstep: at
// every ply the games where the network is to move search for `moveMs` of wall time together,
// then the games where the playout player is to move search for the same time, with the same
// number of leaves per game per iteration.  Openings are seeded random plies from the initial
// position, and each opening is played twice with the colours exchanged.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

struct NetMatchConfig {
  uint32_t games = 400;          // even: games 2i and 2i+1 share opening i
  uint32_t seed = 2016;          // openings
  uint32_t openingMin = 0, openingMax = 6;  // random plies before the players take over
  double moveMs = 100.0;         // wall time of each side's search phase per ply
  uint32_t leafBatch = 4096;     // leaves per iteration over the phase's games
  uint32_t maxPlies = 300;       // games still running after this many plies are draws
  uint32_t playoutSeed = 12345;
  uint32_t netBatch = 4096;      // network batch (max_batch)
  uint32_t progressEvery = 0;    // print a status line to stderr every this many plies (0: never)
  // exploration factor k of each player's tree, sqrt(k ln N / n): the network player's default is
  // the run_ai network player's, the playout player's the reference's (src/mcts.cpp:189-190)
  double netExploration = 64.0;
  double playoutExploration = 2.0;
  // both trees count the chooser's own wins in UCB1 (GameTree::setMoverWins), not the
  // reference's child side to move
  bool moverWins = false;
  int device = 0;
};

struct NetMatchResult {
  uint32_t games = 0, netWins = 0, playoutWins = 0, draws = 0;
  double score = 0, scoreStdErr = 0;  // network's (wins + draws / 2) / games
  uint64_t netLeaves = 0, playoutLeaves = 0;
  uint64_t netIterations = 0, playoutIterations = 0;
  uint64_t plies = 0;
  double seconds = 0;
  // wall time of the search phases by part, [0] network player, [1] playout player: tree
  // selection, leaf scoring (network evaluation or device playouts), tree update
  double selectMs[2] = {0, 0}, scoreMs[2] = {0, 0}, updateMs[2] = {0, 0};
  std::vector<int8_t> outcome;  // per game: 1 network won, -1 playout player won, 0 draw
};

// params: MI_MLP_PARAMS values (the network's weights).
NetMatchResult runNetMatch(const NetMatchConfig &cfg, const float *params);
