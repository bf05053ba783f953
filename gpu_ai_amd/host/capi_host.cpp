// capi_host.cpp -- C entry points of the host layer (include/gpuai/gpuai_host.h).
#include <algorithm>
#include <cstring>
#include <exception>
#include <stdexcept>

#include "gpuai/gpuai_host.h"
#include "mcts.hpp"
#include "multisearch.hpp"
#include "netmatch.hpp"
#include "playout.hpp"
#include "state.hpp"

namespace {

mi_move toMi(const Move &m) {
  mi_move o;
  std::memset(&o, 0, sizeof(o));
  o.from = m.from.square();
  o.to = m.to.square();
  o.jumps = m.jumps;
  o.promoted = m.promoted ? 1 : 0;
  for (unsigned i = 0; i < m.jumps; i++) {
    o.inter[i] = m.intermediate[i].square();
    o.removed[i] = m.removed[i].square();
  }
  return o;
}

uint32_t fmix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}

uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

uint32_t stateKey(const mi_pos &p) {
  return p.p1 ^ rotl(p.p2, 7) ^ rotl(p.kings, 13) ^ ((uint32_t)p.turn << 30) ^
         ((uint32_t)p.counter << 20);
}

}  // namespace

extern "C" {

int gpuai_host_gen_moves(const mi_pos *pos, mi_move *out, uint32_t cap) {
  Move result[MAX_MOVES];
  const State s = State::unpack(*pos);
  const int n = s.genMoves(result);
  if ((uint32_t)n > cap) return -1;
  for (int i = 0; i < n; i++) out[i] = toMi(result[i]);
  return n;
}

int gpuai_host_apply(mi_pos *pos, const mi_move *move) {
  State s = State::unpack(*pos);
  Move m(Loc::ofSquare(move->from), Loc::ofSquare(move->to), move->jumps, move->promoted != 0);
  for (unsigned i = 0; i < move->jumps && i < MAX_MOVE_JUMPS; i++) {
    m.intermediate[i] = Loc::ofSquare(move->inter[i]);
    m.removed[i] = Loc::ofSquare(move->removed[i]);
  }
  for (const Move &legal : s.getMoves()) {
    if (legal == m) {
      s.move(m);
      *pos = s.pack();
      return 0;
    }
  }
  return -1;
}

int gpuai_mcts_trace(const mi_pos *root, uint32_t iterations, uint32_t initial, double scale,
                     uint32_t *trace, uint32_t trace_cap) {
  if (trace_cap < 5u * iterations + 4u) return -1;
  try {
    GameTree tree(State::unpack(*root));
    uint32_t w = 0;
    for (uint32_t it = 0; it < iterations; it++) {
      unsigned n = (unsigned)(tree.getTotalTrials() * scale);
      if (n == 0) n = initial;
      std::vector<State> leaves = tree.select(n);
      std::vector<PlayerId> results(leaves.size());
      uint32_t digest = 0;
      for (size_t i = 0; i < leaves.size(); i++) {
        const uint32_t key = stateKey(leaves[i].pack());
        digest = fmix32(digest ^ key) + 0x9E3779B9u;
        const uint32_t h = fmix32(key ^ (it * 0x9E3779B9u) ^ ((uint32_t)i * 0x85EBCA6Bu));
        results[i] = (PlayerId)((int)(h % 3u) - 1);
      }
      tree.update(results);
      trace[w++] = (uint32_t)leaves.size();
      trace[w++] = digest;
      trace[w++] = tree.getTotalTrials();
      trace[w++] = (uint32_t)tree.getWins(PLAYER_1);
      trace[w++] = (uint32_t)tree.getWins(PLAYER_2);
    }
    for (PlayerId p : {PLAYER_1, PLAYER_2}) {
      Move m = tree.getOptMove(p);
      trace[w++] = m.from.square();
      trace[w++] = m.to.square();
    }
    return (int)w;
  } catch (const std::exception &) {
    return -2;
  }
}

int gpuai_multisearch_run(const gpuai_multisearch_config *cfg, gpuai_multisearch_result *out,
                          char *err, uint32_t err_len) {
  try {
    MultiSearchConfig c;
    c.searches = cfg->searches;
    c.seed = cfg->seed;
    c.playoutSeed = cfg->playout_seed;
    c.minBudget = cfg->min_budget;
    c.maxBudget = cfg->max_budget;
    c.initial = cfg->initial;
    c.scale = cfg->scale;
    c.policy = cfg->policy;
    c.mode = (MultiSearchMode)cfg->mode;
    c.groups = cfg->groups ? cfg->groups : 2;
    c.autotune = cfg->autotune != 0;
    if (c.maxBudget < c.minBudget || c.searches == 0 || cfg->mode > 3)
      throw std::runtime_error("invalid multi-search configuration");
    if (cfg->hook) {
      const gpuai_batch_hook hook = cfg->hook;
      void *user = cfg->hook_user;
      c.hook = [hook, user](const MultiSearchBatch &b) {
        gpuai_batch_record rec;
        std::memset(&rec, 0, sizeof(rec));
        rec.iteration = b.iteration;
        rec.group = b.group;
        rec.seed = b.seed;
        rec.policy = b.policy;
        rec.npos = (uint32_t)b.pos->size();
        rec.playouts = b.winners->size();
        rec.pos = b.pos->data();
        rec.counts = b.counts->data();
        rec.keys = b.keys->data();
        rec.winners = b.winners->data();
        hook(user, &rec);
      };
    }
    MultiSearchResult r = runMultiSearch(c);
    if (cfg->budgets) std::copy(r.budgets.begin(), r.budgets.end(), cfg->budgets);
    if (cfg->root_trials) std::copy(r.rootTrials.begin(), r.rootTrials.end(), cfg->root_trials);
    out->playouts = r.playouts;
    out->iterations = r.iterations;
    out->pad = 0;
    out->wall_seconds = r.wallSeconds;
    out->device_ms = r.deviceMs;
    out->p50_ms = r.p50Ms;
    out->p99_ms = r.p99Ms;
    out->max_ms = r.maxMs;
    out->playouts_per_second = r.playoutsPerSecond;
    out->root_digest = r.rootDigest;
    out->select_ms = r.selectMs;
    out->playout_ms = r.playoutMs;
    out->update_ms = r.updateMs;
    out->tuning = r.tuning;
    return 0;
  } catch (const std::exception &e) {
    if (err && err_len) {
      std::strncpy(err, e.what(), err_len - 1);
      err[err_len - 1] = 0;
    }
    return -1;
  }
}

int gpuai_playout_policy(const char *name) { return name ? playoutPolicyOf(name) : -1; }

int gpuai_driver_run_playouts(const mi_pos *states, uint32_t n, uint32_t policy, uint32_t seed,
                              uint32_t calls, int8_t *winners, char *err, uint32_t err_len) {
  try {
    if ((n && !states) || (n && calls && !winners)) throw std::runtime_error("null argument");
    DeviceMI355XPlayoutDriver driver(policy, 0, seed);
    std::vector<State> v;
    for (uint32_t i = 0; i < n; i++) v.push_back(State::unpack(states[i]));
    for (uint32_t c = 0; c < calls; c++) {
      const std::vector<PlayerId> r = driver.runPlayouts(v);
      if (r.size() != n) throw std::runtime_error("runPlayouts returned the wrong count");
      for (uint32_t i = 0; i < n; i++) winners[(size_t)c * n + i] = (int8_t)r[i];
    }
    return 0;
  } catch (const std::exception &e) {
    if (err && err_len) {
      std::strncpy(err, e.what(), err_len - 1);
      err[err_len - 1] = 0;
    }
    return -1;
  }
}

int gpuai_search_record(const mi_pos *root, uint32_t iterations, uint32_t initial, double scale,
                        uint32_t policy, uint32_t seed, gpuai_batch_hook hook, void *hook_user,
                        char *err, uint32_t err_len) {
  try {
    if (!root) throw std::runtime_error("null root");
    DeviceMI355XPlayoutDriver driver(policy, 0, seed);
    GameTree tree(State::unpack(*root));
    std::vector<mi_pos> packed;
    std::vector<int8_t> w;
    for (uint32_t it = 0; it < iterations; it++) {
      unsigned n = (unsigned)(tree.getTotalTrials() * scale);  // src/player.cpp:137-139
      if (n == 0) n = initial;
      const std::vector<State> leaves = tree.select(n);
      const std::vector<PlayerId> results = driver.runPlayouts(leaves);
      if (hook) {
        packed.resize(leaves.size());
        w.resize(leaves.size());
        for (size_t i = 0; i < leaves.size(); i++) {
          packed[i] = leaves[i].pack();
          w[i] = (int8_t)results[i];
        }
        gpuai_batch_record rec;
        std::memset(&rec, 0, sizeof(rec));
        rec.iteration = it;
        rec.seed = seed;
        rec.policy = policy;
        rec.npos = (uint32_t)leaves.size();
        rec.playouts = leaves.size();
        rec.pos = packed.data();
        rec.winners = w.data();
        hook(hook_user, &rec);
      }
      tree.update(results);
    }
    return 0;
  } catch (const std::exception &e) {
    if (err && err_len) {
      std::strncpy(err, e.what(), err_len - 1);
      err[err_len - 1] = 0;
    }
    return -1;
  }
}

int gpuai_net_match_run(const gpuai_net_match_config *cfg, const float *params,
                        gpuai_net_match_result *out, int8_t *outcome, char *err, uint32_t err_len) {
  try {
    if (!cfg || !params || !out) throw std::runtime_error("null argument");
    NetMatchConfig c;
    c.games = cfg->games;
    c.seed = cfg->seed;
    c.openingMin = cfg->opening_min;
    c.openingMax = cfg->opening_max;
    c.moveMs = cfg->move_ms;
    c.leafBatch = cfg->leaf_batch;
    c.maxPlies = cfg->max_plies;
    c.playoutSeed = cfg->playout_seed;
    c.netBatch = cfg->net_batch;
    c.progressEvery = cfg->progress_every;
    c.netExploration = cfg->net_exploration;
    c.playoutExploration = cfg->playout_exploration;
    c.moverWins = cfg->mover_wins != 0;
    if (c.openingMax < c.openingMin || c.leafBatch == 0 || c.netBatch == 0 || !(c.moveMs >= 0) ||
        !(c.netExploration > 0) || !(c.playoutExploration > 0))
      throw std::runtime_error("invalid match configuration");
    NetMatchResult r = runNetMatch(c, params);
    out->games = r.games;
    out->net_wins = r.netWins;
    out->playout_wins = r.playoutWins;
    out->draws = r.draws;
    out->score = r.score;
    out->score_stderr = r.scoreStdErr;
    out->net_leaves = r.netLeaves;
    out->playout_leaves = r.playoutLeaves;
    out->net_iterations = r.netIterations;
    out->playout_iterations = r.playoutIterations;
    out->plies = r.plies;
    out->seconds = r.seconds;
    for (int s = 0; s < 2; s++) {
      out->select_ms[s] = r.selectMs[s];
      out->score_ms[s] = r.scoreMs[s];
      out->update_ms[s] = r.updateMs[s];
    }
    if (outcome) std::copy(r.outcome.begin(), r.outcome.end(), outcome);
    return 0;
  } catch (const std::exception &e) {
    if (err && err_len) {
      std::strncpy(err, e.what(), err_len - 1);
      err[err_len - 1] = 0;
    }
    return -1;
  }
}

}  // extern "C"
