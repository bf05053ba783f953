// multisearch.cpp -- many concurrent searches sharing one device (see multisearch.hpp).
#include "multisearch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <set>
#include <stdexcept>
#include <tuple>

#include "checkers.h"
#include "mcts.hpp"
#include "playout.hpp"

namespace {

using Clock = std::chrono::steady_clock;

uint64_t mix64(uint64_t h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h;
}

}  // namespace

void multiSearchSetup(const MultiSearchConfig &cfg, std::vector<mi_pos> &roots,
                      std::vector<uint32_t> &budgets) {
  mi_ctx *ctx = nullptr;
  checkMi(mi_create(&ctx, cfg.device, nullptr), nullptr, "mi_create");
  // draw twice as many candidates as needed and keep the first distinct ones
  std::vector<mi_pos> cand((size_t)cfg.searches * 2 + 64);
  int rc = mi_gen_start_positions(ctx, cfg.seed, (uint32_t)cand.size(), 1, 30, 1, cand.data());
  std::string err = rc ? mi_last_error(ctx) : "";
  mi_destroy(ctx);
  if (rc) throw std::runtime_error("mi_gen_start_positions: " + err);
  std::set<std::tuple<uint32_t, uint32_t, uint32_t, uint8_t, uint8_t>> seen;
  roots.clear();
  for (const mi_pos &p : cand) {
    if (roots.size() == cfg.searches) break;
    if (seen.insert(std::make_tuple(p.p1, p.p2, p.kings, p.turn, p.counter)).second) roots.push_back(p);
  }
  if (roots.size() < cfg.searches) throw std::runtime_error("not enough distinct start positions");
  budgets.resize(cfg.searches);
  const uint32_t span = cfg.maxBudget - cfg.minBudget + 1;
  for (uint32_t s = 0; s < cfg.searches; s++) {
    uint32_t w[4];
    ck::philox(0u, 0u, ck::TAG_SEARCH, 0u, cfg.seed, s, w);
    budgets[s] = cfg.minBudget + w[0] % span;
  }
}

namespace {

// The launch shapes for the searches' batches: the context's launch-shape table filled for every
// size an iteration's batch can take -- one search's `initial` trials, doubling up to `maxBatch`
// (every search at its largest iteration) -- tuned on batches of the roots; or the context's
// default shape.  Returns the shape of the largest batch (reported with the results).
mi_tuning tuneSearches(mi_ctx *ctx, const MultiSearchConfig &cfg, const std::vector<mi_pos> &roots,
                       uint64_t maxBatch) {
  mi_tuning t = {};
  if (cfg.autotune) {
    const uint64_t top = std::min<uint64_t>(std::max<uint64_t>(maxBatch, 1), (1u << 31) - 1);
    std::vector<uint32_t> sizes;
    for (uint64_t s = std::max<uint32_t>(cfg.initial, 1u);; s *= 2) {
      sizes.push_back((uint32_t)std::min(s, top));
      if (s >= top) break;
    }
    std::vector<mi_tune_entry> best(sizes.size());
    checkMi(mi_autotune_sizes(ctx, roots.data(), (uint32_t)roots.size(), cfg.policy, sizes.data(),
                              (uint32_t)sizes.size(), 1, best.data()),
            ctx, "mi_autotune_sizes");
    t = best.back().tuning;
  } else {
    checkMi(mi_get_tuning(ctx, &t), ctx, "mi_get_tuning");
  }
  return t;
}

// Latency percentiles, throughput and the digest of every root's statistics.
void finishResult(MultiSearchResult &res, const std::vector<std::unique_ptr<GameTree>> &trees) {
  res.rootTrials.clear();
  for (auto &t : trees) res.rootTrials.push_back(t->getTotalTrials());
  std::vector<double> lat = res.iterationMs;
  std::sort(lat.begin(), lat.end());
  if (!lat.empty()) {
    res.p50Ms = lat[lat.size() / 2];
    res.p99Ms = lat[std::min(lat.size() - 1, (size_t)std::ceil(0.99 * lat.size()) - 1)];
    res.maxMs = lat.back();
  }
  res.playoutsPerSecond = res.wallSeconds > 0 ? res.playouts / res.wallSeconds : 0;
  uint64_t h = 0;
  for (auto &t : trees) {
    h = mix64(h, t->getTotalTrials());
    h = mix64(h, (uint64_t)t->getWins(PLAYER_1));
    h = mix64(h, (uint64_t)t->getWins(PLAYER_2));
  }
  res.rootDigest = h;
}

// Packs the leaves of searches `ids` into one batch: each run of equal consecutive leaves of a
// search becomes one position with a count (as DeviceMI355XPlayoutDriver::runPlayouts does), in
// search order, keyed (s << 16 | r) for run r of search s.  Searches are packed in parallel into
// slots given by a prefix sum of their run counts.  Returns the batch's playouts; `first`
// receives each search's first playout index.
uint64_t packLeaves(const std::vector<uint32_t> &ids, const std::vector<std::vector<State>> &leaves,
                    std::vector<mi_pos> &pos, std::vector<uint32_t> &counts,
                    std::vector<uint32_t> &keys, std::vector<uint64_t> &first) {
  const long m = (long)ids.size();
  std::vector<size_t> runs(ids.size() + 1, 0);
#pragma omp parallel for schedule(dynamic, 8)
  for (long k = 0; k < m; k++) {
    const std::vector<State> &L = leaves[ids[k]];
    size_t r = 0;
    for (size_t i = 0; i < L.size(); i++) r += (i == 0 || !(L[i] == L[i - 1]));
    runs[k + 1] = r;
  }
  first.assign(ids.size() + 1, 0);
  for (long k = 0; k < m; k++) {
    runs[k + 1] += runs[k];
    first[k + 1] = first[k] + leaves[ids[k]].size();
  }
  pos.resize(runs[m]);
  counts.resize(runs[m]);
  keys.resize(runs[m]);
#pragma omp parallel for schedule(dynamic, 8)
  for (long k = 0; k < m; k++) {
    const std::vector<State> &L = leaves[ids[k]];
    const size_t base = runs[k];
    size_t at = base;
    for (size_t i = 0; i < L.size(); i++) {
      if (i > 0 && L[i] == L[i - 1]) {
        counts[at - 1]++;
      } else {
        pos[at] = L[i].pack();
        counts[at] = 1;
        keys[at] = (ids[k] << 16) | (uint32_t)(at - base);  // fewer than 2^16 runs per search
        at++;
      }
    }
  }
  return first[m];
}

// Trials search s hands out this iteration: max(initial, totalTrials * scale), capped by the
// remaining budget and by kMaxTrialsPerIteration.
unsigned trialsFor(const MultiSearchConfig &cfg, const GameTree &tree, uint32_t remaining) {
  unsigned n = (unsigned)(tree.getTotalTrials() * cfg.scale);
  n = std::max<unsigned>(n, cfg.initial);
  n = std::min<unsigned>(n, remaining);
  return std::min<unsigned>(n, kMaxTrialsPerIteration);
}

// The most trials a search with this budget can hand out in one iteration (graph capacity).
uint64_t iterationCap(const MultiSearchConfig &cfg, uint32_t budget) {
  uint64_t n = std::min<uint64_t>(budget, kMaxTrialsPerIteration);
  if (cfg.scale <= 0.0) n = std::min<uint64_t>(n, std::max<uint32_t>(cfg.initial, 1u));
  return n;
}

void callHook(const MultiSearchConfig &cfg, uint32_t iteration, uint32_t group,
              const std::vector<mi_pos> &pos, const std::vector<uint32_t> &counts,
              const std::vector<uint32_t> &keys, const std::vector<int8_t> &winners) {
  if (!cfg.hook) return;
  MultiSearchBatch b;
  b.iteration = iteration;
  b.group = group;
  b.seed = cfg.playoutSeed;
  b.policy = cfg.policy;
  b.pos = &pos;
  b.counts = &counts;
  b.keys = &keys;
  b.winners = &winners;
  cfg.hook(b);
}

// Hands each search of `ids` its leaves' winners (leaf i of search k is playout first[k] + i).
void scatterResults(const std::vector<uint32_t> &ids,
                    const std::vector<std::vector<State>> &leaves,
                    const std::vector<uint64_t> &first, const std::vector<int8_t> &winners,
                    std::vector<std::vector<PlayerId>> &results) {
  const long m = (long)ids.size();
#pragma omp parallel for schedule(dynamic, 8)
  for (long k = 0; k < m; k++) {
    const uint32_t s = ids[k];
    results[s].resize(leaves[s].size());
    for (size_t i = 0; i < leaves[s].size(); i++)
      results[s][i] = (PlayerId)winners[first[k] + i];
  }
}

// Half of the searches with a device context and graph of its own (pipelined mode).
struct SearchGroup {
  std::vector<uint32_t> ids;
  std::unique_ptr<DeviceMI355XPlayoutDriver> driver;
  mi_graph *graph = nullptr;
  std::vector<mi_pos> pos;
  std::vector<uint32_t> counts, keys;
  std::vector<int8_t> winners;
  std::vector<uint64_t> first;  // each search's first playout in the batch
  uint32_t index = 0;           // the group's number
  uint32_t launches = 0;        // the group's searches' iteration (the next batch's epoch)
  bool launched = false;
  ~SearchGroup() {
    if (graph) mi_graph_destroy(graph);
  }
};

double msSince(Clock::time_point a) {
  return std::chrono::duration<double, std::milli>(Clock::now() - a).count();
}

// Pipelined mode: the searches are dealt round-robin into `groups` groups, each with a device
// context (stream) and graph of its own; groups take turns, so that while one group's tree
// selection and update run on the host the other groups' playouts run on the device.  Each search
// still alternates select -> playouts -> update.
MultiSearchResult runPipelined(const MultiSearchConfig &cfg,
                               std::vector<std::unique_ptr<GameTree>> &trees,
                               std::vector<uint32_t> &remaining, const std::vector<mi_pos> &roots) {
  const uint32_t S = (uint32_t)trees.size();
  const uint32_t ngroups = std::max<uint32_t>(1u, std::min<uint32_t>(cfg.groups, S));
  std::vector<std::unique_ptr<SearchGroup>> groups;
  for (uint32_t k = 0; k < ngroups; k++) {
    groups.emplace_back(new SearchGroup());
    SearchGroup &Gk = *groups.back();
    Gk.index = k;
    for (uint32_t s = k; s < S; s += ngroups) Gk.ids.push_back(s);
    // one seed for every group: the stream keys tell the searches apart
    Gk.driver = std::make_unique<DeviceMI355XPlayoutDriver>(cfg.policy, cfg.device,
                                                            cfg.playoutSeed);
  }
  // each group's context gets its own launch-shape table (each batch then runs with its size's
  // shape; a graph keeps the refill policy it was created with)
  MultiSearchResult res;
  for (uint32_t k = 0; k < ngroups; k++) {
    SearchGroup &Gk = *groups[k];
    uint64_t cap = 0;
    for (uint32_t s : Gk.ids) cap += iterationCap(cfg, remaining[s]);
    cap = std::max<uint64_t>(cap, 1);
    const mi_tuning t = tuneSearches(Gk.driver->context(), cfg, roots, cap);
    if (k == 0) res.tuning = t;
    checkMi(mi_graph_create(Gk.driver->context(), (uint32_t)std::min<uint64_t>(cap, 1u << 22),
                            (uint32_t)std::min<uint64_t>(cap, (1u << 31) - 1), cfg.policy,
                            &Gk.graph),
            Gk.driver->context(), "mi_graph_create");
  }
  std::vector<std::vector<State>> leaves(S);
  std::vector<std::vector<PlayerId>> results(S);

  auto selectAndLaunch = [&](SearchGroup &G) {
    const auto t0 = Clock::now();
    const long m = (long)G.ids.size();
#pragma omp parallel for schedule(dynamic, 4)
    for (long i = 0; i < m; i++) {
      const uint32_t s = G.ids[i];
      leaves[s].clear();
      if (remaining[s] == 0) continue;
      trees[s]->selectInto(trialsFor(cfg, *trees[s], remaining[s]), leaves[s]);
    }
    res.selectMs += msSince(t0);
    const auto t1 = Clock::now();
    packLeaves(G.ids, leaves, G.pos, G.counts, G.keys, G.first);
    G.launched = !G.pos.empty();
    if (G.launched)
      checkMi(mi_graph_launch(G.graph, G.pos.data(), (uint32_t)G.pos.size(), G.counts.data(),
                              G.keys.data(), G.driver->seed(), G.launches),
              G.driver->context(), "mi_graph_launch");
    res.playoutMs += msSince(t1);
  };
  auto finishAndUpdate = [&](SearchGroup &G) {
    const auto t0 = Clock::now();
    const uint64_t total = G.first.back();
    G.winners.resize(total);
    mi_stats st;
    checkMi(mi_graph_finish(G.graph, G.winners.data(), &st), G.driver->context(),
            "mi_graph_finish");
    G.launched = false;
    res.deviceMs += st.kernel_ms;
    callHook(cfg, G.launches, G.index, G.pos, G.counts, G.keys, G.winners);
    G.launches++;
    scatterResults(G.ids, leaves, G.first, G.winners, results);
    res.playoutMs += msSince(t0);
    const auto t1 = Clock::now();
    const long m = (long)G.ids.size();
#pragma omp parallel for schedule(dynamic, 4)
    for (long i = 0; i < m; i++) {
      const uint32_t s = G.ids[i];
      if (leaves[s].empty()) continue;
      trees[s]->update(results[s]);
      remaining[s] -= (uint32_t)leaves[s].size();
    }
    res.updateMs += msSince(t1);
    res.playouts += total;
  };
  auto pendingWork = [&](const SearchGroup &G) {
    for (uint32_t s : G.ids)
      if (remaining[s]) return true;
    return false;
  };

  const auto wall0 = Clock::now();
  for (auto &g : groups) selectAndLaunch(*g);
  for (;;) {
    const auto t0 = Clock::now();
    bool busy = false;
    for (auto &g : groups) {
      if (g->launched) finishAndUpdate(*g);
      if (pendingWork(*g)) selectAndLaunch(*g);  // runs while the other groups' replays run
      busy |= g->launched;
    }
    if (!busy) break;
    res.iterations++;
    res.iterationMs.push_back(msSince(t0));
  }
  res.wallSeconds = std::chrono::duration<double>(Clock::now() - wall0).count();
  return res;
}

}  // namespace

MultiSearchResult runMultiSearch(const MultiSearchConfig &cfg) {
  if (cfg.searches == 0 || cfg.searches > 65536)
    throw std::runtime_error("the multi-search takes 1 to 65,536 searches");
  std::vector<mi_pos> roots;
  std::vector<uint32_t> budgets;
  multiSearchSetup(cfg, roots, budgets);
  const uint32_t S = cfg.searches;
  std::vector<std::unique_ptr<GameTree>> trees;
  for (const mi_pos &p : roots) trees.emplace_back(new GameTree(State::unpack(p)));
  std::vector<uint32_t> remaining = budgets;
  uint64_t maxIter = 0;
  for (uint32_t b : budgets) maxIter += b;

  if (cfg.mode == MultiSearchMode::PIPELINED) {
    MultiSearchResult res = runPipelined(cfg, trees, remaining, roots);
    res.budgets = budgets;
    finishResult(res, trees);
    return res;
  }
  DeviceMI355XPlayoutDriver driver(cfg.policy, cfg.device, cfg.playoutSeed);
  // the largest batch: every search's largest possible iteration
  uint64_t cap = 0;
  for (uint32_t b : budgets) cap += iterationCap(cfg, b);
  cap = std::max<uint64_t>(cap, 1);
  const mi_tuning tuning = tuneSearches(driver.context(), cfg, roots,
                                        cfg.mode == MultiSearchMode::PER_SEARCH
                                            ? iterationCap(cfg, cfg.maxBudget)
                                            : cap);
  mi_graph *graph = nullptr;
  if (cfg.mode == MultiSearchMode::GRAPH) {
    checkMi(mi_graph_create(driver.context(), (uint32_t)std::min<uint64_t>(cap, 1u << 22),
                            (uint32_t)std::min<uint64_t>(cap, (1u << 31) - 1), cfg.policy, &graph),
            driver.context(), "mi_graph_create");
  }

  MultiSearchResult res;
  res.tuning = tuning;
  res.budgets = budgets;
  std::vector<std::vector<State>> leaves(S);
  std::vector<std::vector<PlayerId>> results(S);
  std::vector<mi_pos> pos;
  std::vector<uint32_t> counts, keys;
  std::vector<int8_t> winners;
  std::vector<uint64_t> first;
  std::vector<uint32_t> all(S);
  for (uint32_t s = 0; s < S; s++) all[s] = s;
  const auto wall0 = Clock::now();
  for (;;) {
    const auto t0 = Clock::now();
    bool any = false;
    for (uint32_t s = 0; s < S; s++) any |= remaining[s] > 0;
    if (!any) break;
    // every unfinished search takes part in every iteration, so the iteration count is each
    // active search's own iteration: the epoch of its batch in every mode
    const uint32_t iteration = res.iterations;
#pragma omp parallel for schedule(dynamic, 4)
    for (long s = 0; s < (long)S; s++) {
      leaves[s].clear();
      if (remaining[s] == 0) continue;
      trees[s]->selectInto(trialsFor(cfg, *trees[s], remaining[s]), leaves[s]);
    }
    res.selectMs += msSince(t0);
    const auto t1 = Clock::now();
    uint64_t iterPlayouts = 0;
    if (cfg.mode == MultiSearchMode::PER_SEARCH) {
      // one device call per search (the reference's pattern: each search calls its driver)
      for (uint32_t s = 0; s < S; s++) {
        if (leaves[s].empty()) continue;
        const std::vector<uint32_t> one(1, s);
        iterPlayouts += packLeaves(one, leaves, pos, counts, keys, first);
        mi_stats st;
        driver.runKeyed(pos, counts, keys, iteration, winners, &st);
        res.deviceMs += st.kernel_ms;
        callHook(cfg, iteration, s, pos, counts, keys, winners);
        scatterResults(one, leaves, first, winners, results);
      }
    } else {
      iterPlayouts = packLeaves(all, leaves, pos, counts, keys, first);
      mi_stats st;
      if (cfg.mode == MultiSearchMode::GRAPH) {
        winners.resize(iterPlayouts);
        checkMi(mi_graph_run(graph, pos.data(), (uint32_t)pos.size(), counts.data(), keys.data(),
                             driver.seed(), iteration, winners.data(), &st),
                driver.context(), "mi_graph_run");
      } else {
        driver.runKeyed(pos, counts, keys, iteration, winners, &st);
      }
      res.deviceMs += st.kernel_ms;
      callHook(cfg, iteration, 0, pos, counts, keys, winners);
      scatterResults(all, leaves, first, winners, results);
    }
    res.playoutMs += msSince(t1);
    const auto t2 = Clock::now();
#pragma omp parallel for schedule(dynamic, 4)
    for (long s = 0; s < (long)S; s++) {
      if (leaves[s].empty()) continue;
      trees[s]->update(results[s]);
      remaining[s] -= (uint32_t)leaves[s].size();
    }
    res.updateMs += msSince(t2);
    res.playouts += iterPlayouts;
    res.iterations++;
    res.iterationMs.push_back(msSince(t0));
    if (res.iterations > maxIter) throw std::runtime_error("multi-search did not converge");
  }
  res.wallSeconds = std::chrono::duration<double>(Clock::now() - wall0).count();
  if (graph) mi_graph_destroy(graph);
  finishResult(res, trees);
  return res;
}
