// playout.hpp -- the playout-driver interface of gpu_ai (src/playout.hpp:27-33) and the MI355X
// drivers that implement it through the C ABI of include/gpuai/mi_playout.h.
#pragma once

#include <atomic>

#include <memory>
#include <string>
#include <vector>

#include "gpuai/mi_playout.h"
#include "state.hpp"

class PlayoutDriver {
 public:
  virtual ~PlayoutDriver() {}
  // One playout from each state; result i is the winner of the playout from states[i].
  virtual std::vector<PlayerId> runPlayouts(std::vector<State>) = 0;
  virtual std::string getName() const = 0;
};

// Random (src/playout.cpp:17-32) or heuristic (src/heuristicPlayout.cpp:12-48) playouts on an
// MI355X.  Runs of equal consecutive states become one position with a playout count (terminal
// leaves repeat their state, src/mcts.cpp:65-68).  Each call draws from its own Philox epoch, so
// repeated calls on the same states are independent samples; the sequence is reproducible from
// the seed.  Instances are independent (one HIP context each) and may be used from different
// threads; errors throw std::runtime_error.
class DeviceMI355XPlayoutDriver : public PlayoutDriver {
 public:
  explicit DeviceMI355XPlayoutDriver(uint32_t policy = MI_POLICY_RANDOM, int device = 0,
                                     uint32_t seed = 12345, const mi_tuning *tuning = nullptr);
  ~DeviceMI355XPlayoutDriver() override;
  DeviceMI355XPlayoutDriver(const DeviceMI355XPlayoutDriver &) = delete;
  DeviceMI355XPlayoutDriver &operator=(const DeviceMI355XPlayoutDriver &) = delete;

  std::vector<PlayerId> runPlayouts(std::vector<State> states) override;
  std::string getName() const override {
    return policy_ == MI_POLICY_RANDOM      ? "device_mi355x"
           : policy_ == MI_POLICY_HEURISTIC ? "device_mi355x_heuristic_argmax"
                                            : "device_mi355x_heuristic";
  }

  // Grouped form used by the batched searches: positions with playout counts, winners in
  // position-major order.
  void runGrouped(const std::vector<mi_pos> &pos, const std::vector<uint32_t> &counts,
                  std::vector<int8_t> &winners, mi_stats *stats = nullptr);
  // Grouped form with stream keys and an explicit epoch (mi_run_playouts_keyed): position i's
  // playouts draw from Philox key (seed, keys[i]), whatever else shares the launch.
  void runKeyed(const std::vector<mi_pos> &pos, const std::vector<uint32_t> &counts,
                const std::vector<uint32_t> &keys, uint32_t epoch, std::vector<int8_t> &winners,
                mi_stats *stats = nullptr);
  // Fills the context's launch-shape table for these batch sizes (mi_autotune_sizes on seeded
  // positions: k in [0, 30] random plies), so that each batch runs with its size's fastest shape.
  void tune(const std::vector<uint32_t> &sizes);
  mi_ctx *context() const { return ctx_; }
  uint32_t nextEpoch() { return epoch_++; }
  uint32_t seed() const { return seed_; }
  uint32_t policy() const { return policy_; }

 private:
  mi_ctx *ctx_ = nullptr;
  uint32_t policy_;
  uint32_t seed_;
  std::atomic<uint32_t> epoch_{0};  // a player's worker and getMove may both launch
};

// Name -> driver.  The reference's driver names (src/playout.cpp:189-223) all select the MI355X
// engine: host, device_single, device_multiple, device_coarse, hybrid, hybrid_heuristic (a plain
// hybrid driver there, src/playout.cpp:211-213) and optimal run random playouts;
// host_heuristic, device_heuristic and optimal_heuristic run the noisy heuristic
// (MI_POLICY_HEURISTIC_NOISY, the reference's N(0, 0.11) weight noise).  device_mi355x and
// device_mi355x_heuristic name those two directly, device_mi355x_heuristic_argmax the
// noise-free heuristic with ties broken by the ply's word (MI_POLICY_HEURISTIC).  The
// optimal names, whose reference driver picks a driver by batch size (src/playout.cpp:80-166),
// get a driver whose launch-shape table is tuned at construction for batches of 2^6 to 2^20
// playouts (about a second).
// Unknown names throw std::runtime_error("Unknown playout type").
std::unique_ptr<PlayoutDriver> getPlayoutDriver(std::string name);
// The MI_POLICY_* that getPlayoutDriver(name) plays, or -1 for an unknown name.
int playoutPolicyOf(const std::string &name);

// Throws std::runtime_error with the library's message when rc != MI_OK.
void checkMi(int rc, const mi_ctx *ctx, const char *what);
