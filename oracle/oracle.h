/*
 * oracle.h -- CPU restatement of gpu_ai's host playout path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gpu_ai_amd/) includes,
 * links or loads this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * The restatement keeps the reference's data layout (an 8x8 array of
 * {occupied, type, owner} cells, src/state.hpp:71-108,119-122) and its
 * recursive move generator, so that it follows src/state.cu line by line in
 * behaviour while being written independently.  Positions cross this API in
 * the packed 16-byte form shared with the C ABI of the device library
 * (three 32-bit bitboards over the 32 dark squares + turn + draw counter).
 *
 * Square numbering (src/state.cu:173-176): s = 4*row + (col >> 1) with
 * col = 2*(s & 3) + (row even ? 1 : 0); row 0 holds player 1's men at start.
 */
#ifndef GPUAI_ORACLE_H
#define GPUAI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_MOVES 256   /* list capacity; the reference's is 100 (src/state.hpp:19) */
#define OR_MAX_JUMPS 8     /* src/state.hpp:17 */
#define OR_DRAW_MOVES 50   /* src/state.hpp:14 */

#define OR_PLAYER_1 0
#define OR_PLAYER_2 1
#define OR_PLAYER_NONE (-1)

#define OR_POLICY_RANDOM 0
#define OR_POLICY_HEURISTIC 1          /* the heuristic without noise, ties by the ply's word */
#define OR_POLICY_HEURISTIC_NOISY 2    /* the heuristic with the N(0, 0.11) draw per move */
/* Reference harness only (timing): the random choice drawn with glibc rand() % n, as
 * RandomPlayer::getMove does (src/player.cpp:13-16).  Not reproducible; never used for parity. */
#define OR_POLICY_RANDOM_LIBC 9

/* Philox stream tags (third counter word). */
#define OR_TAG_PLAYOUT 0u
#define OR_TAG_START 1u
#define OR_TAG_SEARCH 2u
#define OR_TAG_MLP 3u

/* Packed position: identical layout to mi_pos in include/gpuai/mi_playout.h. */
typedef struct {
  uint32_t p1;      /* bit s set: player 1 piece on square s */
  uint32_t p2;      /* bit s set: player 2 piece on square s */
  uint32_t kings;   /* bit s set: the piece on s is a king */
  uint8_t turn;     /* 0 = player 1 to move, 1 = player 2 */
  uint8_t counter;  /* plies since the last capture (movesSinceLastCapture) */
  uint8_t pad[2];
} or_pos;

/* Canonical move record (S7): squares are 0..31, lists hold `jumps` entries. */
typedef struct {
  uint8_t from, to, jumps, promoted;
  uint8_t inter[OR_MAX_JUMPS];    /* landing square of jump i */
  uint8_t removed[OR_MAX_JUMPS];  /* captured square of jump i */
} or_move;

/* Per-playout trace written by the batch runners. */
typedef struct {
  int8_t winner;     /* 0, 1 or -1 (draw) */
  uint8_t max_n;     /* largest move-list length seen (saturating at 255) */
  uint16_t pad;
  uint32_t plies;    /* moves applied */
  uint32_t hash;     /* move-list hash chain (or_hash_* below) */
  uint32_t captures; /* plies whose list was a capture list */
} or_trace;

/* ---- Philox4x32-10 (Salmon et al., SC'11) ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* One word of the playout stream: key (seed, pos), counter (ply>>2, playout, tag, epoch). */
uint32_t or_stream_word(uint32_t seed, uint32_t pos, uint32_t ply, uint32_t playout,
                        uint32_t tag, uint32_t epoch);

/* ---- game model ---- */
or_pos or_start_position(void);
int or_gen_moves(const or_pos *p, or_move *out, int cap);         /* returns n, -1 if cap too small */
int or_apply(or_pos *p, const or_move *m);                         /* 0 ok, -1 invalid */
int or_is_game_over(const or_pos *p);
int or_winner(const or_pos *p);                                    /* call only when game over */
uint64_t or_perft(const or_pos *p, int depth);

/* ---- heuristic (src/heuristic.cu) ---- */
void or_score_state(const or_pos *p, uint32_t score[2]);
void or_score_move(const or_pos *p, const or_move *m, int32_t score[2]);

/* ---- hashing of move lists (verification mode) ---- */
uint32_t or_fmix32(uint32_t x);
uint32_t or_move_digest(const or_move *m);
uint32_t or_list_digest(const or_move *list, int n);
uint32_t or_ply_record(uint32_t list_digest, int n, int chosen);   /* chosen = -1 at game end */
uint32_t or_hash_step(uint32_t h, uint32_t record);

/* Noisy heuristic: the draw sigma * z of move m (mover order) in a ply of word r, and the
 * standard normal pairs z[2i], z[2i+1] of the words a[i], b[i]. */
float or_heuristic_noise(uint32_t r, uint32_t m);
void or_normal_pairs(const uint32_t *a, const uint32_t *b, uint32_t n, float *z);
/* Choice frequencies of one ply of `policy` from p over `trials` words; returns n (-1: n > cap). */
int or_choice_counts(const or_pos *p, int policy, uint32_t seed, uint32_t trials, uint32_t *counts,
                     int cap);

/* ---- playouts ---- */
int or_playout(or_pos p, uint32_t seed, uint32_t pos_index, uint32_t playout, uint32_t epoch,
               int policy, or_trace *tr);
/* positions x counts[i] playouts, results in playout order (prefix offsets of counts).
 * counts == NULL means `per_pos` playouts for each position. threads <= 0 -> OpenMP default. */
int or_run_playouts(const or_pos *pos, uint32_t npos, const uint32_t *counts, uint32_t per_pos,
                    uint32_t seed, uint32_t epoch, int policy, int8_t *winners, or_trace *traces,
                    int threads);
/* As or_run_playouts with stream keys: position i's playouts use Philox key (seed, keys[i])
 * instead of (seed, i); keys == NULL is or_run_playouts. */
int or_run_playouts_keyed(const or_pos *pos, uint32_t npos, const uint32_t *counts,
                          uint32_t per_pos, const uint32_t *keys, uint32_t seed, uint32_t epoch,
                          int policy, int8_t *winners, or_trace *traces, int threads);

/* ---- start positions ----
 * k = kmin + (r mod (kmax-kmin+1)) random plies from the start position, stopping early at
 * game over (src/driver.cpp:90-98).  With reject_terminal, a finished position is redrawn with
 * the next attempt counter.  Returns -2 for kmax < kmin or a range of 2^32 ply numbers, -1 when
 * reject_terminal found no unfinished position in 4096 attempts. */
int or_gen_start_positions(uint32_t seed, uint32_t n, uint32_t kmin, uint32_t kmax,
                           int reject_terminal, or_pos *out);

#ifdef __cplusplus
}
#endif
#endif
