// multisearch.hpp -- many concurrent MCTS searches sharing one device.
//
// Each search is a GameTree (the reference's batch UCB1 search, src/mcts.cpp) with its own
// playout budget.  Every iteration each unfinished search selects a batch of leaves
// (MCTSPlayer::worker's select, src/player.cpp:137-143); the leaves of all searches are then
// played out either
//   GRAPH      : in one replay of a captured hipGraph (one persistent refill kernel),
//   BATCHED    : in one ordinary mi_run_playouts call, or
//   PER_SEARCH : one mi_run_playouts call per search (the reference's one-driver-call-per-search
//                pattern, the baseline of the comparison),
//   PIPELINED  : the searches in `groups` groups, each on a context (stream) and graph of its
//                own, taking turns so that one group's selection and update on the host overlap
//                the other groups' replays,
// and every search updates its tree with its own results.
#pragma once

#include <stdint.h>

#include <vector>

#include "gpuai/mi_playout.h"

enum class MultiSearchMode { GRAPH = 0, BATCHED = 1, PER_SEARCH = 2, PIPELINED = 3 };

struct MultiSearchConfig {
  uint32_t searches = 256;
  uint32_t seed = 2016;          // start positions and budgets
  uint32_t playoutSeed = 12345;  // playout streams
  uint32_t minBudget = 1000;     // playouts per search, uniform in [minBudget, maxBudget]
  uint32_t maxBudget = 64000;
  uint32_t initial = 256;        // trials per iteration: max(initial, totalTrials * scale),
  double scale = 0.0;            //   capped by the remaining budget
  uint32_t policy = MI_POLICY_RANDOM;
  MultiSearchMode mode = MultiSearchMode::GRAPH;
  uint32_t groups = 2;           // pipelined mode: groups of searches (device streams)
  // launch shape from the autotuner (mi_autotune on every root with `initial` playouts, before
  // the timed search; the shape changes no result), else the library's default
  bool autotune = true;
  int device = 0;
};

struct MultiSearchResult {
  uint64_t playouts = 0;
  uint32_t iterations = 0;
  double wallSeconds = 0;
  double deviceMs = 0;           // summed device time of the playout launches
  double p50Ms = 0, p99Ms = 0, maxMs = 0;  // per-iteration wall latency
  // wall time by phase, summed over iterations: tree selection, playouts (packing, launch,
  // results), tree update
  double selectMs = 0, playoutMs = 0, updateMs = 0;
  double playoutsPerSecond = 0;
  uint64_t rootDigest = 0;       // digest of every root's statistics (determinism checks)
  mi_tuning tuning = {};         // the launch shape the searches ran with
  std::vector<double> iterationMs;
};

// Seeded distinct start positions (1..30 random plies, as the configuration prescribes) and
// budgets 1000 + (r mod 63001) from Philox key (seed, search), counter (0, 0, 2, 0).
void multiSearchSetup(const MultiSearchConfig &cfg, std::vector<mi_pos> &roots,
                      std::vector<uint32_t> &budgets);

MultiSearchResult runMultiSearch(const MultiSearchConfig &cfg);
