/*
 * mi_playout.h -- C ABI of the MI355X (gfx950) Monte Carlo playout engine for 8x8 checkers.
 *
 * Plain pointers and sizes only.  The caller owns every host buffer; a context owns its device
 * memory, its HIP stream and its timing events.  Every entry point returns MI_OK (0) or an
 * error code that mi_strerror() describes; nothing in the library calls exit().
 *
 * Semantics follow the host path of gpu_ai (krame505/gpu_ai):
 *   move generation   src/state.cu:122-180,239-245,255-438 (State::genMoves and helpers)
 *   move application  src/state.cu:57-92 (State::move)
 *   terminal rule     src/state.cpp:16-23 (isGameOver / getWinner)
 *   random playout    src/playout.cpp:17-32 with src/player.cpp:13-16 (rand() % n)
 *   heuristic playout src/heuristicPlayout.cpp:12-48 with src/heuristic.cu:7-49
 * with each random draw taken from a counter-based Philox4x32-10 stream: the word for ply t of
 * playout j of position i is word (t & 3) of Philox(key = (seed, i),
 * counter = (t >> 2, j, 0, epoch)).  A random ply picks move (r mod n) of the reference-ordered
 * list.  A noisy heuristic ply (MI_POLICY_HEURISTIC_NOISY, the reference's) adds to each move's
 * weight (S_mover + d_mover) / (S_opp + d_opp) an N(0, 0.11) draw made from r and the move's
 * index and plays the first largest sum; where all weights are equal it plays move (r mod n),
 * the argmax of iid draws being uniform there.  A plain heuristic ply (MI_POLICY_HEURISTIC)
 * picks T[r mod |T|] of the reference-ordered set T of moves of maximal weight, compared
 * exactly.
 *
 * Squares: s = 4*row + (col >> 1), col = 2*(s & 3) + (row even ? 1 : 0) (src/state.cu:173-176).
 */
#ifndef GPUAI_MI_PLAYOUT_H
#define GPUAI_MI_PLAYOUT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MI_OK 0
#define MI_ERR_INVALID_ARGUMENT 1
#define MI_ERR_NO_DEVICE 2
#define MI_ERR_DEVICE 3
#define MI_ERR_OUT_OF_MEMORY 4
#define MI_ERR_CAPACITY 5      /* a move list exceeded the caller's capacity */
#define MI_ERR_NOT_STAGED 6

/* ---- policies ---- */
#define MI_POLICY_RANDOM 0     /* HostPlayoutDriver, src/playout.cpp:17-32 */
/* The heuristic of src/heuristicPlayout.cpp:12-48 without its noise: the largest weight
 * (S_me + d_me) / (S_op + d_op), ties broken by the ply's word. */
#define MI_POLICY_HEURISTIC 1
/* HostHeuristicPlayoutDriver and DeviceHeuristicPlayoutDriver (src/heuristicPlayout.cpp:12-48,
 * src/heuristicPlayout.cu:50-62): the largest weight plus an N(0, 0.11) draw per move, the
 * draws taken from the ply's word by a Box-Muller transform. */
#define MI_POLICY_HEURISTIC_NOISY 2

/* ---- refill policies of the persistent playout kernel ---- */
#define MI_REFILL_STATIC 0     /* lane-strided assignment, no atomics */
#define MI_REFILL_QUEUE 1      /* wave-level claims, multiples of 16 playouts, from slab counters */

/* Packed position (16 bytes).  Replaces the 776-byte State (src/state.hpp:119-122). */
typedef struct mi_pos {
  uint32_t p1;       /* bit s: a player-1 piece stands on square s */
  uint32_t p2;       /* bit s: a player-2 piece stands on square s */
  uint32_t kings;    /* bit s: the piece on square s is a king */
  uint8_t turn;      /* 0: player 1 to move (PLAYER_1), 1: player 2 (PLAYER_2) */
  uint8_t counter;   /* movesSinceLastCapture */
  uint8_t pad[2];
} mi_pos;

/* Canonical move record (20 bytes).  Replaces Move (src/state.hpp:253-260); equality as
 * Move::operator== (src/state.cu:440-454): only the first `jumps` entries are meaningful. */
typedef struct mi_move {
  uint8_t from, to, jumps, promoted;
  uint8_t inter[8];    /* intermediate[i]: landing square of jump i */
  uint8_t removed[8];  /* removed[i]: square captured by jump i */
} mi_move;

/* Launch parameters of the persistent playout kernel.  Zero fields take the tuned defaults. */
typedef struct mi_tuning {
  uint32_t block_size;      /* threads per workgroup: 64..1024, a multiple of 64 */
  uint32_t blocks_per_cu;   /* resident workgroups per compute unit for the persistent grid */
  uint32_t refill;          /* MI_REFILL_* */
  uint32_t defer_min;       /* a wave plays its parked plies (multi-jump captures; for the noisy
                               heuristic also plies whose moves weigh differently) once this many
                               lanes wait (or all active lanes wait, or 8 iterations pass) */
} mi_tuning;

/* Per-call statistics. */
typedef struct mi_stats {
  uint64_t playouts;       /* playouts run */
  uint64_t plies;          /* moves applied over all playouts */
  float kernel_ms;         /* device time of the playout kernel (HIP events on the ctx stream) */
  float total_ms;          /* host wall time of the whole call */
  uint32_t max_moves;      /* longest move list seen */
  uint32_t grid_blocks;    /* workgroups launched */
  uint32_t block_size;     /* threads per workgroup */
  uint32_t refill;         /* refill policy used */
} mi_stats;

/* One candidate launch shape measured by mi_autotune. */
typedef struct mi_tune_result {
  mi_tuning tuning;
  float kernel_ms;        /* median device time over the repetitions */
  float playouts_per_s;   /* batch size / kernel_ms */
} mi_tune_result;

typedef struct mi_ctx mi_ctx;

/* Creates a context on HIP device `device`.  `tuning` may be NULL (or have zero fields) for the
 * default launch shape: 64-thread blocks, 12 per CU, queue refill, defer_min 16.  No
 * counterpart in the reference, whose drivers allocate on every call
 * (src/singlePlayout.cu:76-84). */
int mi_create(mi_ctx **ctx, int device, const mi_tuning *tuning);
int mi_destroy(mi_ctx *ctx);
const char *mi_strerror(int code);
/* Text of the last error raised on this context (empty if none). */
const char *mi_last_error(const mi_ctx *ctx);
/* Sets the launch shape, for every batch size: the launch-shape table (see mi_autotune) is
 * emptied.  Zero fields take the defaults of mi_create. */
int mi_set_tuning(mi_ctx *ctx, const mi_tuning *tuning);
int mi_get_tuning(const mi_ctx *ctx, mi_tuning *tuning);
/* Device facts: compute units, wave size, clock in kHz. */
int mi_device_info(const mi_ctx *ctx, uint32_t *compute_units, uint32_t *wave_size,
                   uint32_t *clock_khz);

/* Runs playouts and returns one winner per playout (0, 1, or -1 for a draw).
 * Position i contributes counts[i] playouts (or per_pos each when counts == NULL); winners are
 * laid out position-major.  Replaces PlayoutDriver::runPlayouts (src/playout.hpp:31) as
 * implemented by DeviceSinglePlayoutDriver (src/singlePlayout.cu:71-120),
 * DeviceMultiplePlayoutDriver (src/multiplePlayout.cu:53-98), DeviceCoarsePlayoutDriver
 * (src/coarsePlayout.cu:91-163) and DeviceHeuristicPlayoutDriver
 * (src/heuristicPlayout.cu:102-147).  A total of zero playouts returns MI_OK at once
 * (src/singlePlayout.cu:73-75).  `stats` may be NULL. */
int mi_run_playouts(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                    uint32_t per_pos, uint32_t seed, uint32_t epoch, uint32_t policy,
                    int8_t *winners, mi_stats *stats);

/* As mi_run_playouts, with stream keys: the playouts of position i draw from Philox key
 * (seed, keys[i]) instead of (seed, i), so that a playout's stream depends on the caller's key
 * rather than on where its position sits in the batch (batched searches key each leaf by its
 * search and its place in that search's batch, so that any grouping of the searches into
 * launches plays the same games).  No counterpart in the reference.  keys may be NULL
 * (mi_run_playouts). */
int mi_run_playouts_keyed(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                          uint32_t per_pos, const uint32_t *keys, uint32_t seed, uint32_t epoch,
                          uint32_t policy, int8_t *winners, mi_stats *stats);

/* As mi_run_playouts, and also returns per playout the number of plies and the hash chain of
 * every generated move list (verification mode; slower).  `plies` or `hashes` may be NULL. */
int mi_run_playouts_traced(mi_ctx *ctx, const mi_pos *pos, uint32_t npos,
                           const uint32_t *counts, uint32_t per_pos, uint32_t seed,
                           uint32_t epoch, uint32_t policy, int8_t *winners, uint32_t *plies,
                           uint32_t *hashes, mi_stats *stats);

/* Per-position tallies instead of per-playout winners: tallies[3*i + 0] draws,
 * [3*i + 1] player-1 wins, [3*i + 2] player-2 wins.  Same streams as mi_run_playouts. */
int mi_run_playouts_tally(mi_ctx *ctx, const mi_pos *pos, uint32_t npos,
                          const uint32_t *counts, uint32_t per_pos, uint32_t seed,
                          uint32_t epoch, uint32_t policy, uint32_t *tallies, mi_stats *stats);

/* Generates the legal move list of each position on the device, in the reference order of
 * State::genMoves (src/state.cu:239-245).  moves[i*cap .. i*cap + counts[i]) receives the list
 * of position i.  Returns MI_ERR_CAPACITY if a list is longer than cap.  Verification entry
 * point, as genMovesTest (src/genMovesTest.cu:26-100). */
int mi_gen_moves(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, mi_move *moves, uint32_t cap,
                 uint32_t *counts);

/* Seeded start positions, one per index i < n: k = kmin + (r mod (kmax - kmin + 1)) random plies
 * from the initial position (r = word 0 of Philox(key (seed, i), counter (0, a, 1, 1))), each ply
 * choosing move (r_t mod n_t) with r_t = word (t & 3) of Philox(key (seed, i),
 * counter (t >> 2, a, 1, 0)), stopping early when the game is over.  With reject_terminal, a
 * finished position is redrawn with attempt a + 1.  kmin <= kmax, and kmax - kmin + 1 must fit
 * 32 bits.  Follows genRandomStates (src/driver.cpp:75-104), whose shared engine is not
 * reproducible. */
int mi_gen_start_positions(mi_ctx *ctx, uint32_t seed, uint32_t n, uint32_t kmin, uint32_t kmax,
                           uint32_t reject_terminal, mi_pos *out);

/* Autotuner (no counterpart in the reference, whose OptimalPlayoutDriver picks between whole
 * drivers by a runtime model of the batch size, src/playout.cpp:80-166).  Runs the given batch
 * under every candidate launch shape -- workgroup sizes 64..1024, 1..8 resident waves per SIMD
 * (which sets the playouts each lane runs), both refill policies and deferral thresholds 1, 8, 16
 * and 32 -- `reps` times each and takes the one with the smallest median kernel time.
 *
 * The context keeps a launch-shape table: per policy, batch-size bucket (batches of 2^b to
 * 2^(b+1) - 1 playouts) and refill policy, the fastest shape measured.  Every tuning run records
 * its batch's bucket there; mi_run_playouts (and the keyed, traced, tally and staged launches)
 * and graph replays then launch each batch with its bucket's shape (graph replays: the fastest of
 * the graph's refill policy), and batches whose bucket has no entry with the context's tuning.
 * mi_set_tuning empties the table.  The allocation-free entry point always uses the context's
 * tuning (its batch size is on the device).
 *
 * mi_autotune tunes on one batch and also makes the winner the context's tuning.  `results` (may
 * be NULL) receives up to `max_results` candidates; `nresults` the number measured; `best` (may
 * be NULL) the winner. */
int mi_autotune(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                uint32_t per_pos, uint32_t policy, uint32_t reps, mi_tune_result *results,
                uint32_t max_results, uint32_t *nresults, mi_tuning *best);

/* One entry of the launch-shape table. */
typedef struct mi_tune_entry {
  uint32_t policy;      /* MI_POLICY_* */
  uint32_t bucket;      /* batches of 2^bucket .. 2^(bucket+1) - 1 playouts */
  mi_tuning tuning;     /* the fastest shape measured (its refill field names the refill policy) */
  float kernel_ms;      /* its median kernel time on the measured batch */
  uint32_t playouts;    /* size of the measured batch */
} mi_tune_entry;

/* Fills the table for several batch sizes: for each size S, the first min(S, npos) positions
 * with S / min(S, npos) playouts each (S rounded down to a multiple of the count) are tuned as
 * mi_autotune tunes a batch.  The context's tuning is left as it was.  `best` (may be NULL)
 * receives one entry per size: the fastest shape over both refill policies. */
int mi_autotune_sizes(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, uint32_t policy,
                      const uint32_t *sizes, uint32_t nsizes, uint32_t reps, mi_tune_entry *best);

/* The table's entries (policy, bucket, refill order); `count` receives their number, of which
 * up to `cap` are written. */
int mi_get_tuning_table(const mi_ctx *ctx, mi_tune_entry *out, uint32_t cap, uint32_t *count);

/* Graph-replayed batches for search loops (no counterpart in the reference, which launches,
 * copies and frees per call).  mi_graph_create fixes the capacities and builds one hipGraph for
 * the context's stream: a reset kernel, the persistent playout kernel (the launch shape of a full
 * batch, whose refill policy the graph keeps; each replay takes its batch's shape from the
 * launch-shape table, see mi_autotune, and its grid is sized to the batch), and the copy of the
 * winners back to pinned memory.
 * mi_graph_run packs the batch (seed, epoch, positions, playout offsets and, when keys is not
 * NULL, stream keys) into pinned memory, copies it to the device with one asynchronous copy on
 * the stream, replays the graph and returns the winners as mi_run_playouts_keyed would (same
 * streams).  kernel_ms in the stats is the device time of the upload and the replay.  A graph is
 * used by one thread at a time and must be destroyed before its context. */
typedef struct mi_graph mi_graph;
int mi_graph_create(mi_ctx *ctx, uint32_t max_positions, uint32_t max_playouts, uint32_t policy,
                    mi_graph **graph);
int mi_graph_run(mi_graph *graph, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                 const uint32_t *keys, uint32_t seed, uint32_t epoch, int8_t *winners,
                 mi_stats *stats);
/* The two halves of mi_graph_run, so that the caller can work while the replay runs: launch
 * copies the batch into the graph's pinned buffer and enqueues the replay on the context's
 * stream; finish waits for it and returns its winners.  A graph holds one replay in flight;
 * graphs of different contexts run concurrently on their own streams. */
int mi_graph_launch(mi_graph *graph, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
                    const uint32_t *keys, uint32_t seed, uint32_t epoch);
int mi_graph_finish(mi_graph *graph, int8_t *winners, mi_stats *stats);
int mi_graph_destroy(mi_graph *graph);

/* Device-resident batches (inputs stay in HBM between launches).
 * mi_stage uploads positions and playout counts once; mi_launch_staged runs the staged batch
 * and leaves winners on the device; mi_fetch_winners copies them back. */
int mi_stage(mi_ctx *ctx, const mi_pos *pos, uint32_t npos, const uint32_t *counts,
             uint32_t per_pos);
int mi_launch_staged(mi_ctx *ctx, uint32_t seed, uint32_t epoch, uint32_t policy,
                     mi_stats *stats);
int mi_fetch_winners(mi_ctx *ctx, int8_t *winners, uint64_t n);

/* Allocation-free variant of mi_run_playouts, for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops that forbid hipMalloc).  No counterpart in the
 * reference.  Positions, counts and winners are DEVICE pointers owned by the caller: pos holds
 * npos positions; counts (may be NULL: per_pos playouts each) npos counts; winners one byte per
 * playout, position-major, as mi_run_playouts returns them (same streams, same results).  The
 * call enqueues its kernels on `stream` (a hipStream_t; NULL: the context's stream) and returns
 * without waiting; the results are ready when the stream has run them.  It allocates nothing:
 * the kernels' work words, offsets and hand-out order live in a fixed device array of about
 * 1.75 MiB per device, set up by mi_create.  At most MI_NOALLOC_MAX_POSITIONS positions (larger
 * batches return MI_ERR_INVALID_ARGUMENT), fewer than 2^31 playouts.  Positions and counts in
 * device memory are checked on the device: status (a device word, may be NULL) receives MI_OK,
 * or MI_ERR_INVALID_ARGUMENT when a position is malformed or the counts add up to 2^31 or more,
 * in which case nothing is played.  All calls on one device share its scratch array: calls that
 * overlap in time on one device (other streams or threads) race; calls in sequence on one
 * stream do not. */
#define MI_NOALLOC_MAX_POSITIONS 65536u
int mi_run_playouts_noalloc(mi_ctx *ctx, const mi_pos *pos, uint32_t npos,
                            const uint32_t *counts, uint32_t per_pos, uint32_t seed,
                            uint32_t epoch, uint32_t policy, int8_t *winners, uint32_t *status,
                            void *stream);

/* Diagnostic: runs the device's exact quotient/remainder routines on caller data, q[i] =
 * r[i] / n[i] and rem[i] = r[i] % n[i], n[i] >= 1; rem[i] is all ones where the general
 * remainder, or for n[i] < 2^18 the one that picks r mod n on every ply, disagrees with the
 * quotient routine.  No counterpart in the reference, whose draws use the % operator on the host
 * (src/player.cpp:15). */
int mi_check_divmod(mi_ctx *ctx, const uint32_t *r, const uint32_t *n, uint32_t count,
                    uint32_t *q, uint32_t *rem);

/* Diagnostic: copies `bytes` (a multiple of 16) between two device buffers with a plain
 * vectorised copy kernel (k_copy), `repeats` times, and returns the average kernel time.  Used to
 * calibrate the HBM traffic counters (FETCH_SIZE / WRITE_SIZE) against a known byte count. */
int mi_copy_calibration(mi_ctx *ctx, uint64_t bytes, uint32_t repeats, float *avg_ms);

/* Diagnostic (no counterpart in the reference): the device's measured int32 VALU issue rate, the
 * peak of the playout kernels' roofline.  Runs a probe kernel (k_valu_rate: 32 waves per CU, 8
 * per SIMD, eight independent chains of v_add_u32 / v_xor_b32 per lane, exactly 32 VALU
 * instructions per lane per iteration) `repeats` times over `iters` iterations and returns
 * the best rate in lane-operations per second and the average kernel time.  iters in
 * [1, 1<<20], repeats >= 1. */
int mi_valu_rate(mi_ctx *ctx, uint32_t iters, uint32_t repeats, double *lane_ops_per_s,
                 float *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
