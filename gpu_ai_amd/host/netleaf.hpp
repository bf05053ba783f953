// netleaf.hpp -- leaves scored by the value network (include/gpuai/mi_mlp.h) instead of played
// out: the alternative to PlayoutDriver::runPlayouts (src/playout.hpp:31) for GameTree searches.
#pragma once

#include <string>
#include <vector>

#include "gpuai/mi_mlp.h"
#include "state.hpp"

class NetLeafEvaluator {
 public:
  // A network on `ctx`'s device (the context is borrowed and must outlive the evaluator) for
  // batches of up to maxBatch positions; larger leaf sets are scored in several batches.
  NetLeafEvaluator(mi_ctx *ctx, const float *params, uint32_t maxBatch = 4096);
  NetLeafEvaluator(mi_ctx *ctx, const std::string &checkpoint, uint32_t maxBatch = 4096);
  // As above with a context of its own on `device` (players without a playout driver).
  NetLeafEvaluator(int device, const std::string &checkpoint, uint32_t maxBatch = 4096);
  ~NetLeafEvaluator();
  NetLeafEvaluator(const NetLeafEvaluator &) = delete;
  NetLeafEvaluator &operator=(const NetLeafEvaluator &) = delete;

  // Player 1's win probability at each leaf, for GameTree::updateValues: finished games score
  // their result (1, 0, or 1/2 for a draw); the others (1 + v) / 2 for the side to move, where v
  // in [-1, 1] is the network's expected result for that side.
  void p1WinProb(const std::vector<State> &leaves, std::vector<float> &out);
  std::vector<float> p1WinProb(const std::vector<State> &leaves);

 private:
  void init(uint32_t maxBatch);
  void load(const std::string &checkpoint);
  mi_ctx *ctx_;
  bool ownsCtx_ = false;
  mi_mlp *net_ = nullptr;
  uint32_t maxBatch_;
};

// Path of the network checkpoint used by the net-leaf players: $GPUAI_VALUE_NET, else
// "value.ckpt" in the working directory.
std::string valueNetPath();
