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
      tree(newTree(getStartingState())) {}

MCTSPlayer::MCTSPlayer(unsigned initialNumPlayouts, float numPlayoutsScale, double timeout,
                       std::unique_ptr<NetLeafEvaluator> net)
    : initialNumPlayouts(initialNumPlayouts),
      numPlayoutsScale(numPlayoutsScale),
      timeout(timeout),
      net(std::move(net)),
      tree(newTree(getStartingState())) {}

MCTSPlayer::~MCTSPlayer() { stop(); }

std::shared_ptr<GameTree> MCTSPlayer::newTree(const State &state) const {
  auto t = std::make_shared<GameTree>(state);
  t->setParallel(true);
  if (net) t->setExploration(kNetExploration);
  return t;
}

Move MCTSPlayer::getMove(const State &state, bool verbose) {
  {
    std::lock_guard<std::mutex> lock(treeMutex);
    if (state != tree->state) {
      tree = newTree(state);
      updateCancelled = true;
    }
  }
  std::this_thread::sleep_for(std::chrono::duration<double>(timeout));
  std::lock_guard<std::mutex> lock(treeMutex);
  if (workerError) std::rethrow_exception(workerError);  // the search failed on the device
  if (verbose) {
    std::cout << "Tree size: " << tree->getTotalTrials() << std::endl;
    for (int i = 0; i < NUM_PLAYERS; i++)
      std::cout << (PlayerId)i << " score: " << tree->getScore((PlayerId)i) << std::endl;
  }
  if (!tree->isExpanded() || tree->getTotalTrials() == 0) {
    // too little search time for a first update (the worker's batch may be in flight, its root
    // expanded but scoreless): drop that batch and grow the root here
    updateCancelled = true;
    std::vector<State> leaves = tree->select(2 * (unsigned)tree->getMoves().size() + 2);
    if (net) tree->updateValues(net->p1WinProb(leaves));
    else tree->update(playoutDriver->runPlayouts(leaves));
  }
  return tree->getOptMove(state.turn);
}

void MCTSPlayer::move(const Move &m) {
  std::lock_guard<std::mutex> lock(treeMutex);
  tree = tree->move(m);
  tree->setParallel(true);
  updateCancelled = true;
}

// A game starts from the initial position (playGame, src/driver.cpp:177-183).  The reference
// keeps the previous game's tree here, so the opponent's first move of a second game is applied
// to the last game's final position; the tree is reset instead.
void MCTSPlayer::start() {
  if (workerThread.joinable()) {
    {
      std::lock_guard<std::mutex> lock(treeMutex);
      if (running) return;
    }
    workerThread.join();  // a worker that stopped on an error
  }
  std::lock_guard<std::mutex> lock(treeMutex);
  if (running) return;
  tree = newTree(getStartingState());
  workerError = nullptr;
  running = true;
  workerThread = std::thread(&MCTSPlayer::worker, this);
}

void MCTSPlayer::stop() {
  {
    std::lock_guard<std::mutex> lock(treeMutex);
    updateCancelled = true;
    running = false;
  }
  if (workerThread.joinable()) workerThread.join();  // also after the worker stopped on an error
}

SearchStats MCTSPlayer::stats() const {
  std::lock_guard<std::mutex> lock(treeMutex);
  return stats_;
}

// src/player.cpp:134-150: select under the lock, play out without it, update under it unless a
// move arrived meanwhile.
void MCTSPlayer::worker() {
  try {
    search();
  } catch (...) {  // handed to the next getMove instead of terminating the process
    std::lock_guard<std::mutex> lock(treeMutex);
    workerError = std::current_exception();
    running = false;
  }
}

void MCTSPlayer::search() {
  std::unique_lock<std::mutex> lock(treeMutex);
  while (running) {
    unsigned numPlayouts = (unsigned)(tree->getTotalTrials() * numPlayoutsScale);
    if (numPlayouts == 0) numPlayouts = initialNumPlayouts;
    std::shared_ptr<GameTree> t = tree;
    std::vector<State> leaves = t->select(numPlayouts);
    updateCancelled = false;
    lock.unlock();
    std::vector<PlayerId> results;
    std::vector<float> values;
    if (net) net->p1WinProb(leaves, values);
    else results = playoutDriver->runPlayouts(leaves);
    lock.lock();
    if (!updateCancelled && t == tree) {
      if (net) tree->updateValues(values);
      else tree->update(results);
      stats_.iterations++;
      stats_.playouts += leaves.size();
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
  if (name == "mcts_device_mi355x_net")
    return std::make_unique<MCTSPlayer>(4096, 0.05f, timeout >= 0 ? timeout : 7.0,
                                        std::make_unique<NetLeafEvaluator>(0, valueNetPath()));
  if (name == "mcts_device_mi355x_heuristic")
    return mcts(4096, 0.05f, 7, "device_mi355x_heuristic");
  throw std::runtime_error("Unknown player type");
}
