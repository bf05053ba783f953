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
//
// Streams: run r of search s's leaves in the search's iteration i is played with Philox key
// (playoutSeed, s << 16 | r) and epoch i (mi_run_playouts_keyed, keyed graph replays), so a
// search's playouts do not depend on which other searches share its launches: every mode grows
// the same trees.  A search hands out at most 65,536 trials per iteration (the run index's 16
// bits), and there are at most 65,536 searches.
#pragma once

#include <stdint.h>

#include <functional>
#include <vector>

#include "gpuai/mi_playout.h"

// One device batch of the searches, with its winners (handed to MultiSearchConfig::hook after
// the winners arrive and before the trees are updated).
struct MultiSearchBatch {
  uint32_t iteration = 0;  // the searches' iteration, also the batch's epoch
  uint32_t group = 0;      // pipelined mode: the group; per-search mode: the search; else 0
  uint32_t seed = 0, policy = 0;
  const std::vector<mi_pos> *pos = nullptr;
  const std::vector<uint32_t> *counts = nullptr;
  const std::vector<uint32_t> *keys = nullptr;
  const std::vector<int8_t> *winners = nullptr;
};

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
  std::function<void(const MultiSearchBatch &)> hook;  // test hook: every batch (may be empty)
};

constexpr uint32_t kMaxTrialsPerIteration = 65536;  // leaves per search per iteration

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
  std::vector<uint32_t> budgets;     // each search's playout budget
  std::vector<uint32_t> rootTrials;  // each search's root trials at the end
};

// Seeded distinct start positions (1..30 random plies, as the configuration prescribes) and
// budgets 1000 + (r mod 63001) from Philox key (seed, search), counter (0, 0, 2, 0).
void multiSearchSetup(const MultiSearchConfig &cfg, std::vector<mi_pos> &roots,
                      std::vector<uint32_t> &budgets);

MultiSearchResult runMultiSearch(const MultiSearchConfig &cfg);
