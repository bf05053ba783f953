// netmatch.cpp -- value-network leaves against playout leaves (see netmatch.hpp).
#include "netmatch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <stdexcept>

#include "mcts.hpp"
#include "netleaf.hpp"
#include "playout.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Game {
  State state;
  std::shared_ptr<GameTree> tree[2];  // [0] network player's, [1] playout player's
  PlayerId netSide = PLAYER_1;
  uint32_t plies = 0;
  bool done = false;
  int8_t outcome = 0;
};

double msSince(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

// One side's search phase over the games `ids` (whose mover is that side's player): select
// leaves in every tree, score them in one batch, update every tree, until moveMs has passed.
template <class Score, class Update>
void searchPhase(std::vector<Game> &games, const std::vector<uint32_t> &ids, int side,
                 const NetMatchConfig &cfg, Score score, Update update, NetMatchResult &res) {
  if (ids.empty()) return;
  const unsigned per = std::max<unsigned>(2u, cfg.leafBatch / (unsigned)ids.size());
  std::vector<std::vector<State>> sel(ids.size());
  std::vector<State> batch;
  std::vector<size_t> offs(ids.size() + 1);
  const auto t0 = Clock::now();
  do {
    auto t = Clock::now();
#pragma omp parallel for schedule(dynamic, 2)
    for (long i = 0; i < (long)ids.size(); i++) {
      sel[i].clear();
      games[ids[i]].tree[side]->selectInto(per, sel[i]);
    }
    batch.clear();
    for (size_t i = 0; i < ids.size(); i++) {
      offs[i] = batch.size();
      batch.insert(batch.end(), sel[i].begin(), sel[i].end());
    }
    offs[ids.size()] = batch.size();
    res.selectMs[side] += msSince(t);
    t = Clock::now();
    score(batch);
    res.scoreMs[side] += msSince(t);
    t = Clock::now();
    update(offs);
    res.updateMs[side] += msSince(t);
    (side == 0 ? res.netLeaves : res.playoutLeaves) += batch.size();
    (side == 0 ? res.netIterations : res.playoutIterations)++;
  } while (msSince(t0) < cfg.moveMs);
}

}  // namespace

NetMatchResult runNetMatch(const NetMatchConfig &cfg, const float *params) {
  if (cfg.games == 0 || cfg.games % 2) throw std::invalid_argument("games must be even and > 0");
  DeviceMI355XPlayoutDriver driver(MI_POLICY_RANDOM, cfg.device, cfg.playoutSeed);
  mi_ctx *ctx = driver.context();
  // the playout player's batches (leafBatch leaves over the games of a phase, fewer as games
  // end) each run with their size's fastest launch shape
  {
    std::vector<uint32_t> sizes;
    for (uint32_t s = std::max<uint32_t>(cfg.leafBatch, 1u); s >= 16u; s /= 4) sizes.push_back(s);
    driver.tune(sizes);
  }
  NetLeafEvaluator net(ctx, params, cfg.netBatch);

  // openings: one per pair of games
  const uint32_t pairs = cfg.games / 2;
  std::vector<mi_pos> openings(pairs);
  checkMi(mi_gen_start_positions(ctx, cfg.seed, pairs, cfg.openingMin, cfg.openingMax, 1,
                                 openings.data()),
          ctx, "mi_gen_start_positions");
  std::vector<Game> games(cfg.games);
  for (uint32_t g = 0; g < cfg.games; g++) {
    Game &G = games[g];
    G.state = State::unpack(openings[g / 2]);
    G.netSide = (g % 2 == 0) ? PLAYER_1 : PLAYER_2;
    for (int s = 0; s < 2; s++) {
      G.tree[s] = std::make_shared<GameTree>(G.state);
      G.tree[s]->setParallel(false);
    }
    G.tree[0]->setExploration(cfg.netExploration);
    G.tree[1]->setExploration(cfg.playoutExploration);
    G.tree[0]->setMoverWins(cfg.moverWins);
    G.tree[1]->setMoverWins(cfg.moverWins);
  }

  // leaf scorers: player 1's win probability per leaf, or playout winners
  std::vector<float> p1;
  auto netScore = [&](const std::vector<State> &batch) { net.p1WinProb(batch, p1); };
  std::vector<PlayerId> winners;
  auto playoutScore = [&](const std::vector<State> &batch) { winners = driver.runPlayouts(batch); };

  NetMatchResult res;
  res.games = cfg.games;
  const auto wall0 = Clock::now();
  std::vector<uint32_t> ids[2];
  std::vector<Move> chosen(cfg.games);
  for (uint32_t round = 0;; round++) {
    ids[0].clear();
    ids[1].clear();
    for (uint32_t g = 0; g < cfg.games; g++) {
      Game &G = games[g];
      if (G.done) continue;
      if (G.state.isGameOver() || G.plies >= cfg.maxPlies) {
        const PlayerId w = G.state.isGameOver() ? G.state.getWinner() : PLAYER_NONE;
        G.outcome = w == PLAYER_NONE ? 0 : (w == G.netSide ? 1 : -1);
        G.done = true;
        continue;
      }
      ids[G.state.turn == G.netSide ? 0 : 1].push_back(g);
    }
    if (ids[0].empty() && ids[1].empty()) break;
    if (cfg.progressEvery && round % cfg.progressEvery == 0) {
      uint32_t w = 0, l = 0, d = 0;
      for (const Game &G : games)
        if (G.done) (G.outcome > 0 ? w : G.outcome < 0 ? l : d)++;
      std::fprintf(stderr, "[match] ply %u: %zu games running, network %u wins, playouts %u wins, "
                   "%u draws, %.1f s\n", round, ids[0].size() + ids[1].size(), w, l, d,
                   std::chrono::duration<double>(Clock::now() - wall0).count());
      std::fflush(stderr);
    }
    for (int side = 0; side < 2; side++) {
      if (side == 0) {
        searchPhase(games, ids[0], 0, cfg, netScore,
                    [&](const std::vector<size_t> &o) {
#pragma omp parallel for schedule(dynamic, 2)
                      for (long i = 0; i < (long)ids[0].size(); i++)
                        games[ids[0][i]].tree[0]->updateValues(p1.data() + o[i], o[i + 1] - o[i]);
                    },
                    res);
      } else {
        searchPhase(games, ids[1], 1, cfg, playoutScore,
                    [&](const std::vector<size_t> &o) {
#pragma omp parallel for schedule(dynamic, 2)
                      for (long i = 0; i < (long)ids[1].size(); i++)
                        games[ids[1][i]].tree[1]->update(winners.data() + o[i], o[i + 1] - o[i]);
                    },
                    res);
      }
      for (uint32_t g : ids[side]) chosen[g] = games[g].tree[side]->getOptMove(games[g].state.turn);
    }
    // both players observe every move (MCTSPlayer::move, src/player.cpp:111-116)
#pragma omp parallel for schedule(dynamic, 2)
    for (long k = 0; k < (long)(ids[0].size() + ids[1].size()); k++) {
      const uint32_t g = k < (long)ids[0].size() ? ids[0][k] : ids[1][k - ids[0].size()];
      Game &G = games[g];
      G.state.move(chosen[g]);
      for (int s = 0; s < 2; s++) {
        G.tree[s] = G.tree[s]->move(chosen[g]);
        G.tree[s]->setParallel(false);
      }
      G.plies++;
    }
    res.plies += ids[0].size() + ids[1].size();
  }
  res.seconds = std::chrono::duration<double>(Clock::now() - wall0).count();
  res.outcome.resize(cfg.games);
  double sum = 0, sum2 = 0;
  for (uint32_t g = 0; g < cfg.games; g++) {
    const int8_t o = games[g].outcome;
    res.outcome[g] = o;
    if (o > 0) res.netWins++;
    else if (o < 0) res.playoutWins++;
    else res.draws++;
    const double x = o > 0 ? 1.0 : o < 0 ? 0.0 : 0.5;
    sum += x;
    sum2 += x * x;
  }
  res.score = sum / cfg.games;
  const double var = std::max(0.0, sum2 / cfg.games - res.score * res.score);
  res.scoreStdErr = std::sqrt(var / cfg.games);
  return res;
}
