// playout.cpp -- MI355X playout drivers behind PlayoutDriver::runPlayouts.
#include "playout.hpp"

#include <stdexcept>

void checkMi(int rc, const mi_ctx *ctx, const char *what) {
  if (rc == MI_OK) return;
  std::string msg = std::string(what) + ": " + mi_strerror(rc);
  const char *detail = ctx ? mi_last_error(ctx) : "";
  if (detail && *detail) msg += std::string(" (") + detail + ")";
  throw std::runtime_error(msg);
}

DeviceMI355XPlayoutDriver::DeviceMI355XPlayoutDriver(uint32_t policy, int device, uint32_t seed,
                                                     const mi_tuning *tuning)
    : policy_(policy), seed_(seed) {
  if (policy != MI_POLICY_RANDOM && policy != MI_POLICY_HEURISTIC &&
      policy != MI_POLICY_HEURISTIC_NOISY)
    throw std::runtime_error("unknown playout policy");
  checkMi(mi_create(&ctx_, device, tuning), nullptr, "mi_create");
}

DeviceMI355XPlayoutDriver::~DeviceMI355XPlayoutDriver() { mi_destroy(ctx_); }

void DeviceMI355XPlayoutDriver::tune(const std::vector<uint32_t> &sizes) {
  if (sizes.empty()) return;
  std::vector<mi_pos> pos(1024);
  checkMi(mi_gen_start_positions(ctx_, 2016, (uint32_t)pos.size(), 0, 30, 0, pos.data()), ctx_,
          "mi_gen_start_positions");
  checkMi(mi_autotune_sizes(ctx_, pos.data(), (uint32_t)pos.size(), policy_, sizes.data(),
                            (uint32_t)sizes.size(), 1, nullptr),
          ctx_, "mi_autotune_sizes");
}

void DeviceMI355XPlayoutDriver::runGrouped(const std::vector<mi_pos> &pos,
                                           const std::vector<uint32_t> &counts,
                                           std::vector<int8_t> &winners, mi_stats *stats) {
  uint64_t total = 0;
  for (uint32_t c : counts) total += c;
  winners.resize(total);
  checkMi(mi_run_playouts(ctx_, pos.data(), (uint32_t)pos.size(), counts.data(), 0, seed_,
                          nextEpoch(), policy_, winners.data(), stats),
          ctx_, "mi_run_playouts");
}

void DeviceMI355XPlayoutDriver::runKeyed(const std::vector<mi_pos> &pos,
                                         const std::vector<uint32_t> &counts,
                                         const std::vector<uint32_t> &keys, uint32_t epoch,
                                         std::vector<int8_t> &winners, mi_stats *stats) {
  if (keys.size() != pos.size() || counts.size() != pos.size())
    throw std::runtime_error("runKeyed: one count and one key per position");
  uint64_t total = 0;
  for (uint32_t c : counts) total += c;
  winners.resize(total);
  checkMi(mi_run_playouts_keyed(ctx_, pos.data(), (uint32_t)pos.size(), counts.data(), 0,
                                keys.data(), seed_, epoch, policy_, winners.data(), stats),
          ctx_, "mi_run_playouts_keyed");
}

std::vector<PlayerId> DeviceMI355XPlayoutDriver::runPlayouts(std::vector<State> states) {
  std::vector<PlayerId> results(states.size());
  if (states.empty()) return results;  // src/singlePlayout.cu:73-75
  std::vector<mi_pos> pos;
  std::vector<uint32_t> counts;
  for (size_t i = 0; i < states.size(); i++) {
    if (i > 0 && states[i] == states[i - 1]) {
      counts.back()++;
    } else {
      pos.push_back(states[i].pack());
      counts.push_back(1);
    }
  }
  std::vector<int8_t> winners;
  runGrouped(pos, counts, winners);
  for (size_t i = 0; i < states.size(); i++) results[i] = (PlayerId)winners[i];
  return results;
}

int playoutPolicyOf(const std::string &name) {
  // hybrid_heuristic is a plain hybrid driver in the reference (src/playout.cpp:211-213)
  static const char *random_names[] = {"host", "device_single", "device_multiple", "device_coarse",
                                       "hybrid", "hybrid_heuristic", "optimal", "device_mi355x"};
  static const char *noisy_names[] = {"host_heuristic", "device_heuristic", "optimal_heuristic",
                                      "device_mi355x_heuristic"};
  uint32_t policy = ~0u;
  for (const char *n : random_names)
    if (name == n) policy = MI_POLICY_RANDOM;
  for (const char *n : noisy_names)
    if (name == n) policy = MI_POLICY_HEURISTIC_NOISY;
  if (name == "device_mi355x_heuristic_argmax") policy = MI_POLICY_HEURISTIC;
  return policy == ~0u ? -1 : (int)policy;
}

std::unique_ptr<PlayoutDriver> getPlayoutDriver(std::string name) {
  const int policy = playoutPolicyOf(name);
  if (policy < 0) throw std::runtime_error("Unknown playout type");
  auto d = std::make_unique<DeviceMI355XPlayoutDriver>((uint32_t)policy);
  if (name == "optimal" || name == "optimal_heuristic") {
    std::vector<uint32_t> sizes;
    for (uint32_t s = 1u << 6; s <= (1u << 20); s <<= 2) sizes.push_back(s);
    d->tune(sizes);
  }
  return d;
}
