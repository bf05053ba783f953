// run_ai.cpp -- command line of gpu_ai (src/driver.cpp:26-394) on the MI355X engine.
//
//   run_ai -m gen_moves_test -n N          host vs device move lists of N seeded random states
//   run_ai -m playout_test -n N -1 D1 -2 D2  two playout drivers on N seeded random states
//   run_ai -m game_test -n N -1 P1 -2 P2     N games between two players
//   run_ai [-m game] -1 P1 -2 P2             one game, printed
// Options: --timeout S (seconds per MCTS move), --seed X (random states), -h.
// Random states are 1..100 random plies from the start position, stopping at game over
// (src/driver.cpp:76-104), here drawn from the device's seeded generator.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpuai/mi_playout.h"
#include "player.hpp"
#include "playout.hpp"
#include "state.hpp"

namespace {

enum class Mode { GenMovesTest, PlayoutTest, GameTest, Game };

struct Options {
  Mode mode = Mode::Game;
  unsigned n = 0;
  bool nSet = false;
  std::string p1, p2;
  double timeout = -1.0;
  uint32_t seed = 2016;
};

void usage(const char *argv0) {
  std::cout << "usage: " << argv0
            << " [-m gen_moves_test|playout_test|game_test|game] [-n N] [-1 X] [-2 Y]"
               " [--timeout SECONDS] [--seed S]\n"
               "  players: random mcts mcts_host mcts_host_heuristic mcts_device_coarse"
               " mcts_device_multiple mcts_hybrid mcts_optimal mcts_heuristic"
               " mcts_device_mi355x mcts_device_mi355x_heuristic mcts_device_mi355x_net\n"
               "  (mcts_device_mi355x_net scores leaves with the value network in\n"
               "   $GPUAI_VALUE_NET, default ./value.ckpt)\n"
               "  playout drivers: host host_heuristic device_single device_heuristic"
               " device_multiple device_coarse hybrid hybrid_heuristic optimal"
               " optimal_heuristic device_mi355x device_mi355x_heuristic"
               " device_mi355x_heuristic_argmax\n"
               "  (the *_heuristic drivers add N(0, 0.11) noise to each move's weight, as the"
               " reference's do; hybrid_heuristic plays random playouts, as the reference's"
               " factory makes it; device_mi355x_heuristic_argmax is the heuristic without"
               " noise)\n";
}

Mode parseMode(const std::string &s) {
  if (s == "gen_moves_test") return Mode::GenMovesTest;
  if (s == "playout_test") return Mode::PlayoutTest;
  if (s == "game_test") return Mode::GameTest;
  if (s == "game") return Mode::Game;
  throw std::runtime_error("Unknown run mode");
}

std::vector<State> randomStates(unsigned n, uint32_t seed) {
  mi_ctx *ctx = nullptr;
  checkMi(mi_create(&ctx, 0, nullptr), nullptr, "mi_create");
  std::vector<mi_pos> pos(n);
  int rc = mi_gen_start_positions(ctx, seed, n, 1, 100, 0, pos.data());
  std::string err = rc ? mi_last_error(ctx) : "";
  mi_destroy(ctx);
  checkMi(rc, nullptr, ("mi_gen_start_positions " + err).c_str());
  std::vector<State> states;
  for (const mi_pos &p : pos) states.push_back(State::unpack(p));
  return states;
}

// genMovesTests (src/driver.cpp:106-117): the host generator against the device's.
int genMovesTests(unsigned n, uint32_t seed) {
  std::vector<State> states = randomStates(n, seed);
  std::vector<mi_pos> pos;
  for (const State &s : states) pos.push_back(s.pack());
  mi_ctx *ctx = nullptr;
  checkMi(mi_create(&ctx, 0, nullptr), nullptr, "mi_create");
  const uint32_t cap = 128;
  std::vector<mi_move> moves((size_t)n * cap);
  std::vector<uint32_t> counts(n);
  checkMi(mi_gen_moves(ctx, pos.data(), n, moves.data(), cap, counts.data()), ctx, "mi_gen_moves");
  mi_destroy(ctx);
  std::cout << "Testing " << n << " random states have the same moves on the host and device"
            << std::endl;
  for (unsigned i = 0; i < n; i++) {
    std::vector<Move> host = states[i].getMoves();
    bool same = host.size() == counts[i];
    for (size_t k = 0; same && k < host.size(); k++) {
      const mi_move &d = moves[(size_t)i * cap + k];
      Move m(Loc::ofSquare(d.from), Loc::ofSquare(d.to), d.jumps, d.promoted != 0);
      for (unsigned t = 0; t < d.jumps; t++) {
        m.intermediate[t] = Loc::ofSquare(d.inter[t]);
        m.removed[t] = Loc::ofSquare(d.removed[t]);
      }
      same = m == host[k];
    }
    if (!same) {
      std::cout << "Failed on state " << i << "\n" << states[i] << std::endl;
      return 1;
    }
  }
  std::cout << "Passed" << std::endl;
  return 0;
}

std::vector<PlayerId> playoutTest(const std::vector<State> &states, PlayoutDriver &driver) {
  auto t1 = std::chrono::steady_clock::now();
  std::vector<PlayerId> results = driver.runPlayouts(states);
  auto t2 = std::chrono::steady_clock::now();
  int wins[3] = {0, 0, 0};
  for (PlayerId r : results) wins[r == PLAYER_NONE ? 0 : (r == PLAYER_1 ? 1 : 2)]++;
  std::cout << "=== Results ===\nGames drawn: " << wins[0] << "\nPlayer 1 wins: " << wins[1]
            << "\nPlayer 2 wins: " << wins[2] << "\nElapsed time: "
            << std::chrono::duration<double>(t2 - t1).count() << " seconds" << std::endl;
  return results;
}

// playoutTests (src/driver.cpp:149-170)
int playoutTests(unsigned n, uint32_t seed, const std::string &d1, const std::string &d2) {
  std::vector<State> states = randomStates(n, seed);
  std::unique_ptr<PlayoutDriver> drivers[2] = {getPlayoutDriver(d1), getPlayoutDriver(d2)};
  std::vector<PlayerId> results[2];
  for (int i = 0; i < 2; i++) {
    std::cout << "Running test " << i << ": " << drivers[i]->getName() << "..." << std::endl;
    results[i] = playoutTest(states, *drivers[i]);
    std::cout << std::endl;
  }
  unsigned same = 0;
  for (unsigned i = 0; i < n; i++) same += results[0][i] == results[1][i];
  std::cout << "Similar results: " << (n ? 100.0 * same / n : 0.0) << "%" << std::endl;
  return 0;
}

// playGame (src/driver.cpp:177-232)
PlayerId playGame(std::unique_ptr<Player> players[2], bool verbose) {
  State state = getStartingState();
  for (int i = 0; i < 2; i++) players[i]->start();
  while (!state.isGameOver()) {
    if (verbose) std::cout << state << std::endl;
    Move move = players[state.turn]->getMove(state, verbose);
    if (verbose) std::cout << state.turn << " moved " << move << std::endl;
    for (int i = 0; i < 2; i++) players[i]->move(move);
    state.move(move);
  }
  if (verbose) std::cout << state << std::endl;
  for (int i = 0; i < 2; i++) players[i]->stop();
  PlayerId result = state.getWinner();
  if (verbose) {
    if (result == PLAYER_NONE) std::cout << "Draw!" << std::endl;
    else std::cout << result << " won!" << std::endl;
  }
  return result;
}

int gameTests(unsigned n, std::unique_ptr<Player> players[2]) {
  int wins[3] = {0, 0, 0};
  for (unsigned i = 0; i < n; i++) {
    std::cout << "Playing game " << i << "... " << std::flush;
    PlayerId r = playGame(players, false);
    if (r == PLAYER_NONE) std::cout << "Draw!" << std::endl;
    else std::cout << r << " won!" << std::endl;
    wins[r == PLAYER_NONE ? 0 : (r == PLAYER_1 ? 1 : 2)]++;
  }
  std::cout << "=== Results ===\nGames drawn: " << wins[0] << "\nPlayer 1 wins: " << wins[1]
            << "\nPlayer 2 wins: " << wins[2] << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::cout << "run_ai : checkers player with MCTS on the MI355X playout engine" << std::endl;
  Options o;
  try {
    for (int i = 1; i < argc; i++) {
      std::string a = argv[i];
      auto next = [&]() -> std::string {
        if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
        return argv[++i];
      };
      if (a == "-h" || a == "--help") {
        usage(argv[0]);
        return 0;
      } else if (a == "-m" || a == "--mode") {
        o.mode = parseMode(next());
      } else if (a == "-n" || a == "--num_tests") {
        o.n = (unsigned)std::stoul(next());
        o.nSet = true;
      } else if (a == "-1" || a == "--player1") {
        o.p1 = next();
      } else if (a == "-2" || a == "--player2") {
        o.p2 = next();
      } else if (a == "--timeout") {
        o.timeout = std::stod(next());
      } else if (a == "--seed") {
        o.seed = (uint32_t)std::stoul(next());
      } else if (o.p1.empty()) {
        o.p1 = a;  // positional player 1 (src/driver.cpp:291)
      } else if (o.p2.empty()) {
        o.p2 = a;
      } else {
        throw std::runtime_error("unexpected argument " + a);
      }
    }
    if (!o.nSet) o.n = o.mode == Mode::GameTest ? 10 : 100000;
    switch (o.mode) {
      case Mode::GenMovesTest:
        return genMovesTests(o.n, o.seed);
      case Mode::PlayoutTest:
        return playoutTests(o.n, o.seed, o.p1.empty() ? "host" : o.p1,
                            o.p2.empty() ? "device_coarse" : o.p2);
      case Mode::GameTest:
      case Mode::Game: {
        std::unique_ptr<Player> players[2] = {getPlayer(o.p1.empty() ? "random" : o.p1, o.timeout),
                                              getPlayer(o.p2.empty() ? "mcts" : o.p2, o.timeout)};
        std::cout << "Playing " << (o.mode == Mode::Game ? 1u : o.n) << " game(s) with "
                  << players[0]->getName() << " Player 1 and " << players[1]->getName()
                  << " Player 2" << std::endl;
        if (o.mode == Mode::Game) {
          playGame(players, true);
          return 0;
        }
        return gameTests(o.n, players);
      }
    }
  } catch (std::exception &e) {
    std::cout << e.what() << std::endl;
    usage(argv[0]);
    return 1;
  }
  return 0;
}
