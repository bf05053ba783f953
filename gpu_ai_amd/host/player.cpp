// player.cpp -- random and MCTS players (behaviour of src/player.cpp:13-186).
#include "player.hpp"

#include <chrono>
#include <iostream>
#include <stdexcept>

Move RandomPlayer::getMove(const State &state, bool) {
  std::vector<Move> moves = state.getMoves();
  if (moves.empty()) throw std::runtime_error("no legal move");
  return moves[gen_() % moves.size()];
}

MCTSPlayer::MCTSPlayer(unsigned initialNumPlayouts, float numPlayoutsScale, double timeout,
                       std::unique_ptr<PlayoutDriver> playoutDriver)
    : initialNumPlayouts(initialNumPlayouts),
      numPlayoutsScale(numPlayoutsScale),
      timeout(timeout),
      playoutDriver(std::move(playoutDriver)),
      tree(std::make_shared<GameTree>(getStartingState())) {
  tree->setParallel(true);
}

MCTSPlayer::~MCTSPlayer() {
  if (running) stop();
}

Move MCTSPlayer::getMove(const State &state, bool verbose) {
  {
    std::lock_guard<std::mutex> lock(treeMutex);
    if (state != tree->state) {
      tree = std::make_shared<GameTree>(state);
      tree->setParallel(true);
      updateCancelled = true;
    }
  }
  std::this_thread::sleep_for(std::chrono::duration<double>(timeout));
  std::lock_guard<std::mutex> lock(treeMutex);
  if (verbose) {
    std::cout << "Tree size: " << tree->getTotalTrials() << std::endl;
    for (int i = 0; i < NUM_PLAYERS; i++)
      std::cout << (PlayerId)i << " score: " << tree->getScore((PlayerId)i) << std::endl;
  }
  if (!tree->isExpanded()) {  // too little search time to expand the root: grow it now
    std::vector<State> leaves = tree->select(2 * (unsigned)tree->getMoves().size() + 2);
    tree->update(playoutDriver->runPlayouts(leaves));
  }
  return tree->getOptMove(state.turn);
}

void MCTSPlayer::move(const Move &m) {
  std::lock_guard<std::mutex> lock(treeMutex);
  tree = tree->move(m);
  tree->setParallel(true);
  updateCancelled = true;
}

void MCTSPlayer::start() {
  std::lock_guard<std::mutex> lock(treeMutex);
  if (running) return;
  running = true;
  workerThread = std::thread(&MCTSPlayer::worker, this);
}

void MCTSPlayer::stop() {
  {
    std::lock_guard<std::mutex> lock(treeMutex);
    if (!running) return;
    updateCancelled = true;
    running = false;
  }
  workerThread.join();
}

SearchStats MCTSPlayer::stats() const {
  std::lock_guard<std::mutex> lock(treeMutex);
  return stats_;
}

// src/player.cpp:134-150: select under the lock, play out without it, update under it unless a
// move arrived meanwhile.
void MCTSPlayer::worker() {
  std::unique_lock<std::mutex> lock(treeMutex);
  while (running) {
    unsigned numPlayouts = (unsigned)(tree->getTotalTrials() * numPlayoutsScale);
    if (numPlayouts == 0) numPlayouts = initialNumPlayouts;
    std::shared_ptr<GameTree> t = tree;
    std::vector<State> leaves = t->select(numPlayouts);
    updateCancelled = false;
    lock.unlock();
    std::vector<PlayerId> results = playoutDriver->runPlayouts(leaves);
    lock.lock();
    if (!updateCancelled && t == tree) {
      tree->update(results);
      stats_.iterations++;
      stats_.playouts += results.size();
    }
    updateCancelled = false;
  }
}

std::unique_ptr<Player> getPlayer(std::string name, double timeout) {
  auto mcts = [&](unsigned initial, float scale, double t, const char *driver) {
    return std::make_unique<MCTSPlayer>(initial, scale, timeout >= 0 ? timeout : t,
                                        getPlayoutDriver(driver));
  };
  if (name == "random") return std::make_unique<RandomPlayer>();
  if (name == "mcts") return mcts(MCTS_INITIAL_NUM_PLAYOUTS, MCTS_NUM_PLAYOUTS_SCALE,
                                  MCTS_DEFAULT_TIMEOUT, "device_mi355x");
  if (name == "mcts_host") return mcts(50, 0, 7, "device_mi355x");
  if (name == "mcts_host_heuristic") return mcts(50, 0, 7, "device_mi355x_heuristic");
  if (name == "mcts_device_coarse") return mcts(4000, 0.001f, 7, "device_mi355x");
  if (name == "mcts_device_multiple") return mcts(50, 0.02f, 7, "device_mi355x");
  if (name == "mcts_hybrid") return mcts(50, 0.02f, 7, "device_mi355x");
  if (name == "mcts_optimal") return mcts(50, 0.004f, 7, "device_mi355x");
  if (name == "mcts_heuristic") return mcts(50, 0.004f, 7, "device_mi355x_heuristic");
  if (name == "mcts_device_mi355x") return mcts(4096, 0.05f, 7, "device_mi355x");
  if (name == "mcts_device_mi355x_heuristic")
    return mcts(4096, 0.05f, 7, "device_mi355x_heuristic");
  throw std::runtime_error("Unknown player type");
}
