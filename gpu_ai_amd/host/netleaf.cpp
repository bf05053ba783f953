// netleaf.cpp -- leaves scored by the value network (see netleaf.hpp).
#include "netleaf.hpp"

#include <algorithm>
#include <cstdlib>
#include <stdexcept>

#include "playout.hpp"

NetLeafEvaluator::NetLeafEvaluator(mi_ctx *ctx, const float *params, uint32_t maxBatch)
    : ctx_(ctx), maxBatch_(maxBatch) {
  init(maxBatch);
  checkMi(mi_mlp_set_params(net_, params), ctx_, "mi_mlp_set_params");
}

NetLeafEvaluator::NetLeafEvaluator(mi_ctx *ctx, const std::string &checkpoint, uint32_t maxBatch)
    : ctx_(ctx), maxBatch_(maxBatch) {
  init(maxBatch);
  load(checkpoint);
}

NetLeafEvaluator::NetLeafEvaluator(int device, const std::string &checkpoint, uint32_t maxBatch)
    : ctx_(nullptr), ownsCtx_(true), maxBatch_(maxBatch) {
  checkMi(mi_create(&ctx_, device, nullptr), nullptr, "mi_create");
  try {
    init(maxBatch);
    load(checkpoint);
  } catch (...) {
    mi_destroy(ctx_);
    throw;
  }
}

void NetLeafEvaluator::load(const std::string &checkpoint) {
  if (mi_mlp_load(net_, checkpoint.c_str()) != MI_OK) {
    const std::string why = mi_last_error(ctx_);
    mi_mlp_destroy(net_);
    net_ = nullptr;
    throw std::runtime_error("cannot load the value network " + checkpoint + ": " + why +
                             " (write one with `python -m gpu_ai_amd.value --out PATH` and point "
                             "GPUAI_VALUE_NET at it)");
  }
}

void NetLeafEvaluator::init(uint32_t maxBatch) {
  if (!ctx_ || maxBatch == 0) throw std::invalid_argument("NetLeafEvaluator: no context or batch");
  checkMi(mi_mlp_create(ctx_, maxBatch, 1u, &net_), ctx_, "mi_mlp_create");
}

NetLeafEvaluator::~NetLeafEvaluator() {
  mi_mlp_destroy(net_);
  if (ownsCtx_) mi_destroy(ctx_);
}

void NetLeafEvaluator::p1WinProb(const std::vector<State> &leaves, std::vector<float> &out) {
  // host buffers are local, so calls from several threads only meet in the context's lock
  const size_t count = leaves.size();
  std::vector<mi_pos> all(count), packed;
  std::vector<uint8_t> over(count);
  std::vector<uint32_t> index;
  out.assign(count, 0.5f);
  // finished games are scored by their result, the rest packed for the network (the test and
  // the packing are most of the host time of a call, so they run in parallel over the leaves)
#pragma omp parallel for schedule(static) if (count >= 1024)
  for (long i = 0; i < (long)count; i++) {
    const State &s = leaves[i];
    over[i] = s.isGameOver() ? 1 : 0;
    if (over[i]) {
      const PlayerId w = s.getWinner();
      out[i] = w == PLAYER_1 ? 1.f : w == PLAYER_2 ? 0.f : 0.5f;
    } else {
      all[i] = s.pack();
    }
  }
  packed.reserve(count);
  index.reserve(count);
  for (size_t i = 0; i < count; i++)
    if (!over[i]) {
      packed.push_back(all[i]);
      index.push_back((uint32_t)i);
    }
  std::vector<float> values(packed.size());
  for (size_t a = 0; a < packed.size(); a += maxBatch_) {
    const uint32_t n = (uint32_t)std::min<size_t>(maxBatch_, packed.size() - a);
    checkMi(mi_mlp_forward(net_, packed.data() + a, n, values.data() + a), ctx_, "mi_mlp_forward");
  }
#pragma omp parallel for schedule(static) if (index.size() >= 1024)
  for (long j = 0; j < (long)index.size(); j++) {
    const float v = values[j];
    out[index[j]] = leaves[index[j]].turn == PLAYER_1 ? 0.5f * (1.f + v) : 0.5f * (1.f - v);
  }
}

std::vector<float> NetLeafEvaluator::p1WinProb(const std::vector<State> &leaves) {
  std::vector<float> out;
  p1WinProb(leaves, out);
  return out;
}

std::string valueNetPath() {
  const char *p = std::getenv("GPUAI_VALUE_NET");
  return (p && *p) ? std::string(p) : std::string("value.ckpt");
}
