/*
 * gpuai_host.h -- C entry points of the C++ host layer (libgpuai_host.so).
 *
 * The host layer is the drop-in caller side: State/Move (src/state.hpp), PlayoutDriver
 * (src/playout.hpp:27-33), GameTree (src/mcts.hpp:12-64) and the players (src/player.hpp).
 * These functions expose parts of it to bindings and tests; positions and moves use the
 * layouts of mi_playout.h.
 */
#ifndef GPUAI_HOST_H
#define GPUAI_HOST_H

#include <stdint.h>

#include "gpuai/mi_playout.h"

#ifdef __cplusplus
extern "C" {
#endif

/* State::genMoves on the host (tree expansion uses it); returns the count, -1 if > cap.
 * Replaces State::genMoves (src/state.cu:239-245) for the callers. */
int gpuai_host_gen_moves(const mi_pos *pos, mi_move *out, uint32_t cap);

/* State::move on the host (src/state.cu:57-92); returns 0, or -1 for an illegal move. */
int gpuai_host_apply(mi_pos *pos, const mi_move *move);

/* Grows a GameTree from `root` for `iterations` rounds of select / update, the trial count per
 * round following MCTSPlayer::worker (src/player.cpp:137-139: totalTrials * scale, or `initial`
 * when that is 0).  Each trial's winner is a fixed function of (round, trial index, position)
 * instead of a playout: fmix32(p1 ^ rot(p2,7) ^ rot(kings,13) ^ turn<<30 ^ counter<<20 ^
 * round*0x9E3779B9 ^ trial*0x85EBCA6B) mod 3, minus 1.  Writes 5 words per round (trials
 * handed out, a digest of the selected positions in order, root totalTrials, root wins of player 1
 * and of player 2) followed by the from/to squares of getOptMove for player 1 and player 2.
 * Returns the number of words written, or -1 if trace_cap is too small.  Test hook: the same
 * function over the reference's GameTree lives in the oracle. */
int gpuai_mcts_trace(const mi_pos *root, uint32_t iterations, uint32_t initial, double scale,
                     uint32_t *trace, uint32_t trace_cap);

/* One device batch of playouts as a search loop handed it over, with the winners that came
 * back (test hook: the batch can be replayed on another implementation).  Run r of search s is
 * keyed (s << 16 | r); counts[i] playouts of pos[i] with stream key keys[i] under `seed` and
 * epoch `iteration`; winners position-major. */
typedef struct gpuai_batch_record {
  uint32_t iteration;  /* the searches' iteration = the batch's epoch */
  uint32_t group;      /* pipelined mode: the group; per-search mode: the search; else 0 */
  uint32_t seed, policy, npos, pad;
  uint64_t playouts;
  const mi_pos *pos;
  const uint32_t *counts;
  const uint32_t *keys;
  const int8_t *winners;
} gpuai_batch_record;
typedef void (*gpuai_batch_hook)(void *user, const gpuai_batch_record *record);

/* Batched multi-search: concurrent MCTS searches sharing the device (multisearch.hpp).  The
 * playouts of run r of search s's leaves in the search's iteration i draw from Philox key
 * (playout_seed, s << 16 | r) and epoch i, so every mode grows the same trees. */
typedef struct gpuai_multisearch_config {
  uint32_t searches;      /* 256 */
  uint32_t seed;          /* start positions and budgets (2016) */
  uint32_t playout_seed;  /* playout streams (12345) */
  uint32_t min_budget;    /* 1000 */
  uint32_t max_budget;    /* 64000 */
  uint32_t initial;       /* trials per iteration: max(initial, totalTrials * scale) */
  double scale;
  uint32_t policy;        /* MI_POLICY_* */
  uint32_t mode;          /* 0 graph replay, 1 one plain launch, 2 one launch per search,
                             3 graph replays of `groups` groups overlapped with the host's work */
  uint32_t groups;        /* pipelined mode: groups of searches, one stream each (2) */
  uint32_t autotune;      /* 1: launch shape from mi_autotune before the search, 0: default */
  gpuai_batch_hook hook;  /* called with every device batch (may be NULL) */
  void *hook_user;
  uint32_t *budgets;      /* out (may be NULL): each search's playout budget */
  uint32_t *root_trials;  /* out (may be NULL): each search's root trials at the end */
} gpuai_multisearch_config;

typedef struct gpuai_multisearch_result {
  uint64_t playouts;
  uint32_t iterations;
  uint32_t pad;
  double wall_seconds, device_ms, p50_ms, p99_ms, max_ms, playouts_per_second;
  uint64_t root_digest;
  double select_ms, playout_ms, update_ms;  /* wall time by phase, summed over iterations */
  mi_tuning tuning;                          /* the launch shape the searches ran with */
} gpuai_multisearch_result;

/* Runs the multi-search on device 0; returns 0, or -1 with a message in `err` (err_len bytes). */
int gpuai_multisearch_run(const gpuai_multisearch_config *cfg, gpuai_multisearch_result *out,
                          char *err, uint32_t err_len);

/* Test hooks of the drop-in driver (DeviceMI355XPlayoutDriver, playout.hpp).
 * gpuai_driver_run_playouts: `calls` calls of runPlayouts (src/playout.hpp:31) on the same n
 * states; winners[c * n + i] = result i of call c (call c draws epoch c).
 * gpuai_search_record: one search as MCTSPlayer::worker runs it (src/player.cpp:134-150): a
 * GameTree grown from `root` for `iterations` rounds of select (max(initial, totalTrials * scale)
 * trials), runPlayouts and update; `hook` receives each round's leaves (states in trial order,
 * ungrouped, in record->pos with npos = trials and counts = keys = NULL) and the winners the
 * driver returned.  Both return 0, or -1 with a message in `err`. */
/* The MI_POLICY_* of the driver that getPlayoutDriver(name) (src/playout.cpp:189-223) returns,
 * or -1 for a name it rejects; needs no device. */
int gpuai_playout_policy(const char *name);
int gpuai_driver_run_playouts(const mi_pos *states, uint32_t n, uint32_t policy, uint32_t seed,
                              uint32_t calls, int8_t *winners, char *err, uint32_t err_len);
int gpuai_search_record(const mi_pos *root, uint32_t iterations, uint32_t initial, double scale,
                        uint32_t policy, uint32_t seed, gpuai_batch_hook hook, void *hook_user,
                        char *err, uint32_t err_len);

/* Match of value-network leaves against playout leaves (netmatch.hpp): two MCTS players that
 * differ only in how a leaf is scored, in lockstep games at equal wall time per move.  No
 * counterpart in the reference, whose players always play leaves out (src/player.cpp:135-150). */
typedef struct gpuai_net_match_config {
  uint32_t games;          /* even: each opening is played with both colour assignments (400) */
  uint32_t seed;           /* openings (2016) */
  uint32_t opening_min;    /* random plies before the players take over (0 .. 6) */
  uint32_t opening_max;
  double move_ms;          /* wall time of each side's search phase per ply (100) */
  uint32_t leaf_batch;     /* leaves per iteration over the phase's games (4096) */
  uint32_t max_plies;      /* games still running after this many plies are draws (300) */
  uint32_t playout_seed;   /* (12345) */
  uint32_t net_batch;      /* network batch (4096) */
  uint32_t progress_every; /* status line on stderr every this many plies (0: none) */
  double net_exploration;  /* k of the network player's sqrt(k ln N / n) (64) */
  double playout_exploration; /* k of the playout player's sqrt(k ln N / n) (the reference's 2,
                               * src/mcts.cpp:189-190) */
  uint32_t mover_wins;     /* 0: both trees count the child's side to move in UCB1's win
                            * fraction, as the reference does (src/mcts.cpp:189); 1: the
                            * chooser's own wins (GameTree::setMoverWins) */
} gpuai_net_match_config;

typedef struct gpuai_net_match_result {
  uint32_t games, net_wins, playout_wins, draws;
  double score, score_stderr;   /* network's (wins + draws / 2) / games */
  uint64_t net_leaves, playout_leaves, net_iterations, playout_iterations, plies;
  double seconds;
  /* wall time of the search phases by part, [0] network player, [1] playout player: tree
   * selection, leaf scoring (network evaluation or device playouts), tree update */
  double select_ms[2], score_ms[2], update_ms[2];
} gpuai_net_match_result;

/* params: MI_MLP_PARAMS network weights; outcome (may be NULL): per game 1 network won,
 * -1 playout player won, 0 draw.  Returns 0, or -1 with a message in `err`. */
int gpuai_net_match_run(const gpuai_net_match_config *cfg, const float *params,
                        gpuai_net_match_result *out, int8_t *outcome, char *err, uint32_t err_len);

#ifdef __cplusplus
}
#endif
#endif
