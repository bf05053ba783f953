// player.hpp -- players of gpu_ai (src/player.hpp:20-88): a random player and an MCTS player
// whose background worker grows the tree with batches of playouts from a PlayoutDriver while the
// game waits for its move.
#pragma once

#include <exception>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>

#include "mcts.hpp"
#include "netleaf.hpp"
#include "playout.hpp"

#define MCTS_INITIAL_NUM_PLAYOUTS 50
#define MCTS_NUM_PLAYOUTS_SCALE 0.003
#define MCTS_DEFAULT_TIMEOUT 10  // seconds

class Player {
 public:
  virtual ~Player() {}
  virtual std::string getName() const = 0;
  virtual Move getMove(const State &, bool verbose = true) = 0;
  virtual void start() {}
  virtual void stop() {}
  virtual void move(const Move &) {}
};

// Uniform choice among the legal moves (src/player.cpp:13-16, whose rand() is unseeded; this
// one draws from a seeded engine).
class RandomPlayer : public Player {
 public:
  explicit RandomPlayer(uint64_t seed = 2016) : gen_(seed) {}
  std::string getName() const override { return "random"; }
  Move getMove(const State &state, bool verbose = true) override;

 private:
  std::mt19937_64 gen_;
};

// Search statistics of one MCTSPlayer.
struct SearchStats {
  uint64_t iterations = 0;
  uint64_t playouts = 0;
};

class MCTSPlayer : public Player {
 public:
  // timeout: seconds of search per move (fractions allowed; the reference sleeps whole seconds,
  // src/player.cpp:99)
  MCTSPlayer(unsigned initialNumPlayouts, float numPlayoutsScale, double timeout,
             std::unique_ptr<PlayoutDriver> playoutDriver);
  // Leaves scored by a value network instead of played out (no reference counterpart); its tree
  // explores with kNetExploration.
  MCTSPlayer(unsigned initialNumPlayouts, float numPlayoutsScale, double timeout,
             std::unique_ptr<NetLeafEvaluator> net);
  // Exploration factor k of the network player's UCB1 term sqrt(k ln N / n): its leaf values
  // are already averages, and at equal time per move a far broader search than the reference's
  // k = 2 wins (400-game matches against the playout player, 250 ms per move: k = 2 about 0.53,
  // 8 0.64, 32 0.69, 64 0.71, 128 0.71).
  static constexpr double kNetExploration = 64.0;
  ~MCTSPlayer() override;

  std::string getName() const override {
    return net ? std::string("mcts_device_mi355x_net") : "mcts_" + playoutDriver->getName();
  }
  Move getMove(const State &state, bool verbose = true) override;
  void move(const Move &) override;
  void start() override;
  void stop() override;
  SearchStats stats() const;

 private:
  void worker();
  void search();
  std::shared_ptr<GameTree> newTree(const State &state) const;

  const unsigned initialNumPlayouts;
  const float numPlayoutsScale;
  const double timeout;
  const std::unique_ptr<PlayoutDriver> playoutDriver;
  const std::unique_ptr<NetLeafEvaluator> net;
  std::shared_ptr<GameTree> tree;
  mutable std::mutex treeMutex;
  std::thread workerThread;
  bool updateCancelled = false;
  bool running = false;
  std::exception_ptr workerError;
  SearchStats stats_;
};

// Name -> player.  Reference names (src/player.cpp:152-186) keep their (initial playouts, scale,
// timeout) and run on the MI355X drivers; new names: mcts_device_mi355x,
// mcts_device_mi355x_heuristic, and mcts_device_mi355x_net (leaves scored by the value network
// of the checkpoint valueNetPath() names).  `timeout` < 0 keeps each name's own timeout.
std::unique_ptr<Player> getPlayer(std::string name, double timeout = -1.0);
