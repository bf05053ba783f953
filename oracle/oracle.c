/*
 * oracle.c -- CPU restatement of gpu_ai's host playout path (test infrastructure only).
 *
 * Each function cites the reference file:line whose behaviour it restates.  Where the
 * reference uses a random source that cannot be reproduced (glibc rand(), shared
 * std::default_random_engine, cuRAND), this restatement draws from a counter-based
 * Philox4x32-10 stream instead, one 32-bit word per ply:
 *   playout ply t : key (seed, position), counter (t >> 2, playout, tag 0, epoch), word t & 3
 *   start ply t   : key (seed, position), counter (t >> 2, attempt,  tag 1, 0),     word t & 3
 *   start length  : key (seed, position), counter (0,      attempt,  tag 1, 1),     word 0
 */
#include "oracle.h"

/* OR_MUTATION (off unless defined) plants one known mistake, so that tests can show that the
 * golden checks detect it:
 *   1 men try the (f,+1) jump before (f,-1)      2 kings may land again on their first landing
 *   3 draw at 49 plies instead of 50              4 men keep their type on the far row in captures
 *   5 heuristic ties always take the first move   6 Philox uses 9 rounds
 *   7 the noisy heuristic's sigma is 0.22          8 its noise follows reference, not mover, order
 *   9 it draws on plies whose weights are all equal */
#ifndef OR_MUTATION
#define OR_MUTATION 0
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------ */
/* Philox4x32-10, from Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As Easy as 1, 2,  */
/* 3", SC'11: ten rounds of two 32x32->64 multiplies with a Weyl-sequence key schedule.        */
/* ------------------------------------------------------------------------------------------ */

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < (OR_MUTATION == 6 ? 9 : 10); round++) {
    uint64_t a = (uint64_t)0xD2511F53u * x0;
    uint64_t b = (uint64_t)0xCD9E8D57u * x2;
    uint32_t y0 = (uint32_t)(b >> 32) ^ x1 ^ k0;
    uint32_t y1 = (uint32_t)b;
    uint32_t y2 = (uint32_t)(a >> 32) ^ x3 ^ k1;
    uint32_t y3 = (uint32_t)a;
    x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

uint32_t or_stream_word(uint32_t seed, uint32_t pos, uint32_t ply, uint32_t playout,
                        uint32_t tag, uint32_t epoch) {
  uint32_t ctr[4] = {ply >> 2, playout, tag, epoch};
  uint32_t key[2] = {seed, pos};
  uint32_t out[4];
  or_philox4x32_10(ctr, key, out);
  return out[ply & 3];
}

/* ------------------------------------------------------------------------------------------ */
/* Board model in the reference layout (src/state.hpp:21-31,71-122,253-260).                   */
/* ------------------------------------------------------------------------------------------ */

enum { MAN = 0, KING = 1 };               /* PieceType CHECKER / CHECKER_KING */

typedef struct { uint8_t occupied, type, owner; } cell;     /* BoardItem */
typedef struct { cell b[8][8]; int turn; unsigned since; } board_state;  /* State */
typedef struct { int row, col; } loc;                        /* Loc (signed here) */
typedef struct {
  loc from, to;
  loc removed[OR_MAX_JUMPS], inter[OR_MAX_JUMPS];
  int jumps, promoted;
} board_move;                                                /* Move */

/* Internal list capacity: the DFS may write one slot past the last complete move. */
#define LIST_CAP (OR_MAX_MOVES + 64)

static int loc_on_board(loc l) { return l.row >= 0 && l.row < 8 && l.col >= 0 && l.col < 8; }
static int loc_eq(loc a, loc b) { return a.row == b.row && a.col == b.col; }
static int sq_of(loc l) { return 4 * l.row + (l.col >> 1); }
static loc loc_of(int s) {
  loc l;
  l.row = s >> 2;
  l.col = 2 * (s & 3) + ((l.row & 1) == 0 ? 1 : 0);
  return l;
}
static const cell *at(const board_state *s, loc l) { return &s->b[l.row][l.col]; }

/* src/state.cu:46-55 */
static int next_turn(int turn) { return turn == OR_PLAYER_1 ? OR_PLAYER_2 : OR_PLAYER_1; }

static void unpack(const or_pos *p, board_state *s) {
  memset(s, 0, sizeof(*s));
  for (int q = 0; q < 32; q++) {
    loc l = loc_of(q);
    uint32_t bit = 1u << q;
    if (p->p1 & bit) {
      s->b[l.row][l.col].occupied = 1;
      s->b[l.row][l.col].owner = OR_PLAYER_1;
    } else if (p->p2 & bit) {
      s->b[l.row][l.col].occupied = 1;
      s->b[l.row][l.col].owner = OR_PLAYER_2;
    }
    if ((p->kings & bit) && ((p->p1 | p->p2) & bit)) s->b[l.row][l.col].type = KING;
  }
  s->turn = p->turn ? OR_PLAYER_2 : OR_PLAYER_1;
  s->since = p->counter;
}

static or_pos pack(const board_state *s) {
  or_pos p;
  memset(&p, 0, sizeof(p));
  for (int q = 0; q < 32; q++) {
    const cell *c = at(s, loc_of(q));
    if (!c->occupied) continue;
    if (c->owner == OR_PLAYER_1) p.p1 |= 1u << q; else p.p2 |= 1u << q;
    if (c->type == KING) p.kings |= 1u << q;
  }
  p.turn = (uint8_t)s->turn;
  p.counter = (uint8_t)(s->since > 255 ? 255 : s->since);
  return p;
}

static void move_export(const board_move *m, or_move *o) {
  memset(o, 0, sizeof(*o));
  o->from = (uint8_t)sq_of(m->from);
  o->to = (uint8_t)sq_of(m->to);
  o->jumps = (uint8_t)m->jumps;
  o->promoted = (uint8_t)(m->promoted ? 1 : 0);
  for (int i = 0; i < m->jumps; i++) {
    o->inter[i] = (uint8_t)sq_of(m->inter[i]);
    o->removed[i] = (uint8_t)sq_of(m->removed[i]);
  }
}

static void move_import(const or_move *o, board_move *m) {
  memset(m, 0, sizeof(*m));
  m->from = loc_of(o->from);
  m->to = loc_of(o->to);
  m->jumps = o->jumps;
  m->promoted = o->promoted;
  for (int i = 0; i < o->jumps && i < OR_MAX_JUMPS; i++) {
    m->inter[i] = loc_of(o->inter[i]);
    m->removed[i] = loc_of(o->removed[i]);
  }
}

static board_move move_new(loc from, loc to, int jumps, int promoted) {
  board_move m;
  memset(&m, 0, sizeof(m));
  m.from = from; m.to = to; m.jumps = jumps; m.promoted = promoted;
  return m;
}

/* src/state.cu:456-462: record a landing square and the square jumped on the way. */
static void add_jump(board_move *m, loc land) {
  if (m->jumps >= OR_MAX_JUMPS) abort();        /* the reference asserts (state.cu:457) */
  m->inter[m->jumps] = land;
  m->removed[m->jumps].row = (land.row - m->to.row) / 2 + m->to.row;
  m->removed[m->jumps].col = (land.col - m->to.col) / 2 + m->to.col;
  m->jumps++;
  m->to = land;
}

/* src/state.cu:122-142: the jumped square must hold an opponent piece and the landing square
 * must be empty on the unmodified board; kings also refuse earlier landing squares. */
static int jump_ok(const board_state *s, const board_move *m, loc jumped, loc land, int no_cycles) {
  if (!loc_on_board(land) || !loc_on_board(jumped)) return 0;
  const cell *j = at(s, jumped);
  if (!j->occupied || (int)j->owner == s->turn || at(s, land)->occupied) return 0;
  if (no_cycles)
    for (int i = m->jumps - 1; i >= (OR_MUTATION == 2 ? 1 : 0); i--)
      if (loc_eq(land, m->inter[i])) return 0;
  return 1;
}

/* src/state.cu:283-340: depth-first capture sequences of a man, left branch (col - 1) first,
 * emitting a sequence when no further jump exists; a leaf on the far row is promoted. */
static int captures_man(const board_state *s, loc here, int owner, board_move *res, int count,
                        int first) {
  board_move prev = res[count];
  int dr = owner == OR_PLAYER_1 ? 1 : -1;
  loc jump_l = {here.row + dr, here.col - 1}, land_l = {here.row + 2 * dr, here.col - 2};
  loc jump_r = {here.row + dr, here.col + 1}, land_r = {here.row + 2 * dr, here.col + 2};
  if (OR_MUTATION == 1) {
    loc t = jump_l; jump_l = jump_r; jump_r = t;
    t = land_l; land_l = land_r; land_r = t;
  }
  int ok_l = jump_ok(s, &res[count], jump_l, land_l, 0);
  int ok_r = jump_ok(s, &res[count], jump_r, land_r, 0);
  if (!ok_l && !ok_r && !first) {
    if (OR_MUTATION != 4) {
      if (s->turn == OR_PLAYER_1 && res[count].to.row == 7) res[count].promoted = 1;
      if (s->turn == OR_PLAYER_2 && res[count].to.row == 0) res[count].promoted = 1;
    }
    return count + 1;
  }
  if (ok_l) {
    if (count >= LIST_CAP - 1) abort();
    res[count] = prev;
    add_jump(&res[count], land_l);
    count = captures_man(s, land_l, owner, res, count, 0);
  }
  if (ok_r) {
    if (count >= LIST_CAP - 1) abort();
    res[count] = prev;
    add_jump(&res[count], land_r);
    count = captures_man(s, land_r, owner, res, count, 0);
  }
  return count;
}

/* src/state.cu:388-420: depth-first capture sequences of a king in direction order
 * (+1,+1), (+1,-1), (-1,+1), (-1,-1); validity of all four is decided before descending. */
static int captures_king(const board_state *s, loc here, board_move *res, int count, int first) {
  static const int dr[4] = {1, 1, -1, -1};
  static const int dc[4] = {1, -1, 1, -1};
  board_move prev = res[count];
  loc land[4];
  int ok[4], any = 0;
  for (int i = 0; i < 4; i++) {
    loc jumped = {here.row + dr[i], here.col + dc[i]};
    land[i].row = here.row + 2 * dr[i];
    land[i].col = here.col + 2 * dc[i];
    ok[i] = jump_ok(s, &prev, jumped, land[i], 1) && !(OR_MUTATION == 2 && prev.jumps >= 7);
    any |= ok[i];
  }
  if (!first && !any) return count + 1;
  for (int i = 0; i < 4; i++) {
    if (!ok[i]) continue;
    if (count >= LIST_CAP - 1) abort();
    res[count] = prev;
    add_jump(&res[count], land[i]);
    count = captures_king(s, land[i], res, count, 0);
  }
  return count;
}

/* src/state.cu:255-281: kings try all four directions, men the two forward ones, in the
 * order (+1,+1), (+1,-1), (-1,+1), (-1,-1); a man reaching the far row is promoted. */
static int direct_moves(const board_state *s, loc here, board_move *res) {
  static const int dc[4] = {1, -1, 1, -1};
  static const int dr[4] = {1, 1, -1, -1};
  const cell *c = at(s, here);
  int count = 0, lo = 0, hi = 4;
  if (c->type != KING) {
    lo = c->owner == OR_PLAYER_1 ? 0 : 2;
    hi = c->owner == OR_PLAYER_1 ? 2 : 4;
  }
  for (int i = lo; i < hi; i++) {
    loc to = {here.row + dr[i], here.col + dc[i]};
    if (!loc_on_board(to) || at(s, to)->occupied) continue;
    int promoted = c->type != KING && to.row == (c->owner == OR_PLAYER_1 ? 7 : 0);
    res[count++] = move_new(here, to, 0, promoted);
  }
  return count;
}

/* src/state.cu:144-180: squares in index order (row-major over dark squares), only the pieces
 * of the side to move. */
static int gen_type(const board_state *s, board_move *res, int captures) {
  int n = 0;
  for (int row = 0; row < 8; row++) {
    for (int col = 1 - (row % 2); col < 8; col += 2) {
      loc here = {row, col};
      const cell *c = at(s, here);
      if (!c->occupied || (int)c->owner != s->turn) continue;
      if (n >= OR_MAX_MOVES) abort();
      if (captures) {
        res[n] = move_new(here, here, 0, 0);
        if (c->type == MAN)
          n = captures_man(s, here, c->owner, res, n, 1);
        else
          n = captures_king(s, here, res, n, 1);
      } else {
        n += direct_moves(s, here, &res[n]);
      }
    }
  }
  return n;
}

/* src/state.cu:239-245: captures are mandatory; direct moves only when no capture exists. */
static int gen(const board_state *s, board_move *res) {
  int n = gen_type(s, res, 1);
  if (n == 0) n = gen_type(s, res, 0);
  if (n > OR_MAX_MOVES) abort();
  return n;
}

/* src/state.cu:57-92 */
static void apply(board_state *s, const board_move *m) {
  for (int i = 0; i < m->jumps; i++) s->b[m->removed[i].row][m->removed[i].col].occupied = 0;
  cell mover = *at(s, m->from);
  cell *dst = &s->b[m->to.row][m->to.col];
  dst->occupied = 1;
  dst->type = m->promoted ? KING : mover.type;
  dst->owner = mover.owner;
  s->b[m->from.row][m->from.col].occupied = 0;
  s->turn = next_turn(s->turn);
  if (m->jumps > 0) s->since = 0; else s->since++;
}

/* src/state.cpp:16-23 */
#define DRAW_AT (OR_MUTATION == 3 ? OR_DRAW_MOVES - 1 : OR_DRAW_MOVES)
static int game_over_with(const board_state *s, int n) { return n == 0 || s->since >= DRAW_AT; }
static int winner_of(const board_state *s) {
  return s->since < DRAW_AT ? next_turn(s->turn) : OR_PLAYER_NONE;
}

/* ---- public game-model API ---- */

/* src/state.cpp:25-40 */
or_pos or_start_position(void) {
  or_pos p;
  memset(&p, 0, sizeof(p));
  p.p1 = 0x00000FFFu;   /* squares 0..11: rows 0-2 */
  p.p2 = 0xFFF00000u;   /* squares 20..31: rows 5-7 */
  return p;
}

int or_gen_moves(const or_pos *p, or_move *out, int cap) {
  board_state s;
  board_move list[LIST_CAP];
  unpack(p, &s);
  int n = gen(&s, list);
  if (n > cap) return -1;
  for (int i = 0; i < n; i++) move_export(&list[i], &out[i]);
  return n;
}

int or_apply(or_pos *p, const or_move *m) {
  board_state s;
  board_move bm;
  unpack(p, &s);
  move_import(m, &bm);
  if (!loc_on_board(bm.from) || !at(&s, bm.from)->occupied || (int)at(&s, bm.from)->owner != s.turn)
    return -1;
  apply(&s, &bm);
  *p = pack(&s);
  return 0;
}

int or_is_game_over(const or_pos *p) {
  board_state s;
  board_move list[LIST_CAP];
  unpack(p, &s);
  return game_over_with(&s, gen(&s, list));
}

int or_winner(const or_pos *p) {
  board_state s;
  unpack(p, &s);
  return winner_of(&s);
}

static uint64_t perft_rec(const board_state *s, int depth) {
  if (depth == 0) return 1;
  board_move list[LIST_CAP];
  int n = gen(s, list);
  if (depth == 1) return (uint64_t)n;
  uint64_t total = 0;
  for (int i = 0; i < n; i++) {
    board_state t = *s;
    apply(&t, &list[i]);
    total += perft_rec(&t, depth - 1);
  }
  return total;
}

uint64_t or_perft(const or_pos *p, int depth) {
  board_state s;
  unpack(p, &s);
  return perft_rec(&s, depth);
}

/* ---- heuristic: src/heuristic.hpp:7-8, src/heuristic.cu:20-49 ---- */

static unsigned piece_value(int type) { return type == KING ? 4u : 1u; }

static void score_state(const board_state *s, uint32_t score[2]) {
  score[0] = score[1] = 0;
  for (int row = 0; row < 8; row++)
    for (int col = 1 - (row % 2); col < 8; col += 2)
      if (s->b[row][col].occupied) score[s->b[row][col].owner] += piece_value(s->b[row][col].type);
}

static void score_move(const board_state *s, const board_move *m, int32_t score[2]) {
  score[0] = score[1] = 0;
  if (m->promoted) score[s->turn] += 4 - 1;
  for (int i = 0; i < m->jumps; i++)
    score[next_turn(s->turn)] -= (int32_t)piece_value(at(s, m->removed[i])->type);
}

void or_score_state(const or_pos *p, uint32_t score[2]) {
  board_state s;
  unpack(p, &s);
  score_state(&s, score);
}

void or_score_move(const or_pos *p, const or_move *m, int32_t score[2]) {
  board_state s;
  board_move bm;
  unpack(p, &s);
  move_import(m, &bm);
  score_move(&s, &bm, score);
}

/* ---- move-list hashing (verification mode, defined by this project) ---- */

uint32_t or_fmix32(uint32_t x) {   /* MurmurHash3 finaliser */
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}

uint32_t or_move_digest(const or_move *m) {
  uint32_t d = or_fmix32(0x6A09E667u ^ ((uint32_t)m->from | ((uint32_t)m->to << 5) |
                                        ((uint32_t)m->jumps << 10) | ((uint32_t)m->promoted << 14)));
  for (int j = 0; j < m->jumps; j++)
    d = or_fmix32(d ^ ((uint32_t)m->inter[j] | ((uint32_t)m->removed[j] << 5) |
                       ((uint32_t)(j + 1) << 10)));
  return d;
}

uint32_t or_list_digest(const or_move *list, int n) {
  uint32_t acc = 0;
  for (int i = 0; i < n; i++) acc += or_fmix32(or_move_digest(&list[i]) + (uint32_t)i * 0x9E3779B9u);
  return acc;
}

uint32_t or_ply_record(uint32_t list_digest, int n, int chosen) {
  uint32_t c = chosen < 0 ? 0xFFFFu : (uint32_t)chosen;
  return or_fmix32(list_digest ^ or_fmix32((uint32_t)n | (c << 16)));
}

uint32_t or_hash_step(uint32_t h, uint32_t record) { return or_fmix32(h ^ record) + 0x9E3779B9u; }

static uint32_t list_digest_internal(const board_move *list, int n) {
  uint32_t acc = 0;
  for (int i = 0; i < n; i++) {
    or_move o;
    move_export(&list[i], &o);
    acc += or_fmix32(or_move_digest(&o) + (uint32_t)i * 0x9E3779B9u);
  }
  return acc;
}

/* ---- playouts ---- */

/* Heuristic choice, src/heuristicPlayout.cpp:20-42 without the N(0, 0.11) noise: the weight
 * (S_mover + d_mover) / (S_opp + d_opp) is compared exactly by cross-multiplication (a zero
 * denominator is +infinity; the numerator is at least 1), and the tie set T of maximal moves is
 * broken by T[r mod |T|]. */
static int heuristic_choice(const board_state *s, const board_move *list, int n,
                            const uint32_t score[2], uint32_t r, int32_t chosen_delta[2]) {
  int64_t num[OR_MAX_MOVES], den[OR_MAX_MOVES];
  int32_t delta[OR_MAX_MOVES][2];
  int me = s->turn, op = next_turn(s->turn);
  int best = 0;
  for (int i = 0; i < n; i++) {
    score_move(s, &list[i], delta[i]);
    num[i] = (int64_t)score[me] + delta[i][me];
    den[i] = (int64_t)score[op] + delta[i][op];
    if (num[i] * den[best] > num[best] * den[i]) best = i;
  }
  int ties[OR_MAX_MOVES], nt = 0;
  for (int i = 0; i < n; i++)
    if (num[i] * den[best] == num[best] * den[i]) ties[nt++] = i;
  int pick = OR_MUTATION == 5 ? ties[0] : ties[r % (uint32_t)nt];
  chosen_delta[0] = delta[pick][0];
  chosen_delta[1] = delta[pick][1];
  return pick;
}

/* ---- the noisy heuristic: src/heuristicPlayout.cpp:16,26-36, src/heuristic.hpp:5 ---- */

static float f32_from_bits(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t f32_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/* Two standard normal draws from the words a, b by the Box-Muller transform (Box and Muller,
 * Ann. Math. Statist. 29 (1958) 610-611), in single precision with every step rounded
 * (x86-64 SSE arithmetic, no contraction):
 *   radius = sqrt(-2 ln u), u = ((a >> 8) + 1) 2^-24;  angle = q pi/2 + x, q = b >> 30,
 *   x = (((b >> 6) mod 2^24) 2^-24 - 1/2) pi/2.
 * ln u = (e - 24) ln 2 + 2 atanh((m - 1)/(m + 1)) for (a >> 8) + 1 = m 2^e, m in (sqrt(1/2),
 * sqrt(2)], the atanh series to s^9; sin x and cos x by Taylor series to x^9 and x^10. */
static void normal_pair(uint32_t a, uint32_t b, float *z0, float *z1) {
  const float uf = (float)((a >> 8) + 1u);
  const uint32_t ub = f32_bits(uf);
  int e = (int)(ub >> 23) - 127;
  float m = f32_from_bits((ub & 0x007FFFFFu) | 0x3F800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  const float lnm =
      s * (2.0f + s2 * (0.666666667f + s2 * (0.4f + s2 * (0.285714286f + s2 * 0.222222222f))));
  const float lnu = (float)(e - 24) * 0.693147181f + lnm;
  const float rad = sqrtf(-2.0f * lnu);
  const float x = ((float)((b >> 6) & 0x00FFFFFFu) * 5.96046448e-8f - 0.5f) * 1.57079633f;
  const float x2 = x * x;
  const float sn = x * (1.0f + x2 * (-0.166666667f + x2 * (8.33333333e-3f +
                                     x2 * (-1.98412698e-4f + x2 * 2.75573192e-6f))));
  const float cs = 1.0f + x2 * (-0.5f + x2 * (4.16666667e-2f + x2 * (-1.38888889e-3f +
                                 x2 * (2.48015873e-5f + x2 * -2.75573192e-7f))));
  const uint32_t q = b >> 30;
  const float c = q == 0u ? cs : q == 1u ? -sn : q == 2u ? -cs : sn;
  const float t = q == 0u ? sn : q == 1u ? cs : q == 2u ? -sn : -cs;
  *z0 = rad * c;
  *z1 = rad * t;
}

#define NOISE_SIGMA (OR_MUTATION == 7 ? 0.22f : 0.11f)

/* The noise of move m in mover order (player 2's order is the reverse of the reference's) for
 * the ply word r: moves 2j and 2j+1 take the pair of fmix(r + (2j+1) K), fmix(r + (2j+2) K),
 * K = 0x9E3779B9, times sigma. */
float or_heuristic_noise(uint32_t r, uint32_t m) {
  const uint32_t j = m >> 1;
  float z0, z1;
  normal_pair(or_fmix32(r + (2u * j + 1u) * 0x9E3779B9u), or_fmix32(r + (2u * j + 2u) * 0x9E3779B9u),
              &z0, &z1);
  return ((m & 1u) ? z1 : z0) * NOISE_SIGMA;
}

void or_normal_pairs(const uint32_t *a, const uint32_t *b, uint32_t n, float *z) {
  for (uint32_t i = 0; i < n; i++) normal_pair(a[i], b[i], &z[2 * i], &z[2 * i + 1]);
}

/* Noisy choice, src/heuristicPlayout.cpp:26-36: weight (float)(S_me + d_me) / (S_op + d_op)
 * (src/heuristic.cu:44-49) plus the move's draw; the first move of the largest sum in reference
 * order (strict >, starting from -infinity).  When all n weights are equal the draws alone
 * decide, and the argmax of iid draws is uniform: the ply then plays index r mod n without
 * drawing (index 0 when the common weight is +inf, which no draw changes). */
static int noisy_choice(const board_state *s, const board_move *list, int n, const uint32_t score[2],
                        uint32_t r, int32_t chosen_delta[2]) {
  int me = s->turn, op = next_turn(s->turn);
  float w[OR_MAX_MOVES];
  int32_t d[2];
  int single = 1;
  if (n <= 0) return 0;
  for (int i = 0; i < n; i++) {
    score_move(s, &list[i], d);
    const uint32_t num = score[me] + (uint32_t)d[me];
    const uint32_t den = score[op] + (uint32_t)d[op];
    w[i] = (float)num / (float)den;
    if (w[i] != w[0]) single = 0;
  }
  int pick = 0;
  if (single && OR_MUTATION != 9) {
    pick = isinf(w[0]) ? 0 : (int)(r % (uint32_t)n);
  } else {
    float best = -INFINITY;
    for (int i = 0; i < n; i++) {
      const uint32_t m =
          (OR_MUTATION == 8 || me == OR_PLAYER_1) ? (uint32_t)i : (uint32_t)(n - 1 - i);
      const float v = w[i] + or_heuristic_noise(r, m);
      if (v > best) {
        best = v;
        pick = i;
      }
    }
  }
  score_move(s, &list[pick], chosen_delta);
  return pick;
}

/* Choice frequencies of one ply of `policy` from p: trial t plays the ply with word
 * or_stream_word(seed, 0, t, 0, OR_TAG_PLAYOUT, 0); counts[i] = times reference index i was
 * chosen.  Returns n, or -1 if n > cap. */
int or_choice_counts(const or_pos *p, int policy, uint32_t seed, uint32_t trials, uint32_t *counts,
                     int cap) {
  board_state s;
  board_move list[LIST_CAP];
  uint32_t score[2];
  unpack(p, &s);
  const int n = gen(&s, list);
  if (n > cap) return -1;
  score_state(&s, score);
  for (int i = 0; i < n; i++) counts[i] = 0;
  if (n == 0) return 0;
  for (uint32_t t = 0; t < trials; t++) {
    const uint32_t r = or_stream_word(seed, 0, t, 0, OR_TAG_PLAYOUT, 0);
    int32_t d[2];
    int k;
    if (policy == OR_POLICY_HEURISTIC_NOISY) k = noisy_choice(&s, list, n, score, r, d);
    else if (policy == OR_POLICY_HEURISTIC) k = heuristic_choice(&s, list, n, score, r, d);
    else k = (int)(r % (uint32_t)n);
    counts[k]++;
  }
  return n;
}

/* src/playout.cpp:17-32 (random) and src/heuristicPlayout.cpp:12-48 (heuristic): play until
 * isGameOver (src/state.cpp:16-18), then report getWinner (src/state.cpp:20-23). */
int or_playout(or_pos p, uint32_t seed, uint32_t pos_index, uint32_t playout, uint32_t epoch,
               int policy, or_trace *tr) {
  board_state s;
  board_move list[LIST_CAP];
  uint32_t score[2] = {0, 0};
  uint32_t h = 0, plies = 0, captures = 0;
  int max_n = 0;
  unpack(&p, &s);
  if (policy == OR_POLICY_HEURISTIC || policy == OR_POLICY_HEURISTIC_NOISY) score_state(&s, score);
  uint32_t words[4] = {0, 0, 0, 0};
  for (uint32_t ply = 0;; ply++) {
    int n = gen(&s, list);
    if (n > max_n) max_n = n;
    uint32_t L = tr ? list_digest_internal(list, n) : 0;
    if (game_over_with(&s, n)) {
      if (tr) h = or_hash_step(h, or_ply_record(L, n, -1));
      break;
    }
    if ((ply & 3) == 0) {
      uint32_t ctr[4] = {ply >> 2, playout, OR_TAG_PLAYOUT, epoch};
      uint32_t key[2] = {seed, pos_index};
      or_philox4x32_10(ctr, key, words);
    }
    uint32_t r = words[ply & 3];
    int k;
    if (policy == OR_POLICY_HEURISTIC) {
      int32_t d[2];
      k = heuristic_choice(&s, list, n, score, r, d);
      score[0] += (uint32_t)d[0];
      score[1] += (uint32_t)d[1];
    } else if (policy == OR_POLICY_HEURISTIC_NOISY) {
      int32_t d[2];
      k = noisy_choice(&s, list, n, score, r, d);
      score[0] += (uint32_t)d[0];
      score[1] += (uint32_t)d[1];
    } else {
      k = (int)(r % (uint32_t)n);
    }
    if (list[0].jumps > 0) captures++;
    if (tr) h = or_hash_step(h, or_ply_record(L, n, k));
    apply(&s, &list[k]);
    plies++;
  }
  int w = winner_of(&s);
  if (tr) {
    tr->winner = (int8_t)w;
    tr->max_n = (uint8_t)(max_n > 255 ? 255 : max_n);
    tr->pad = 0;
    tr->plies = plies;
    tr->hash = h;
    tr->captures = captures;
  }
  return w;
}

int or_run_playouts(const or_pos *pos, uint32_t npos, const uint32_t *counts, uint32_t per_pos,
                    uint32_t seed, uint32_t epoch, int policy, int8_t *winners, or_trace *traces,
                    int threads) {
  return or_run_playouts_keyed(pos, npos, counts, per_pos, NULL, seed, epoch, policy, winners,
                               traces, threads);
}

int or_run_playouts_keyed(const or_pos *pos, uint32_t npos, const uint32_t *counts,
                          uint32_t per_pos, const uint32_t *keys, uint32_t seed, uint32_t epoch,
                          int policy, int8_t *winners, or_trace *traces, int threads) {
  uint64_t *offs = (uint64_t *)malloc(sizeof(uint64_t) * (npos + 1));
  if (!offs) return -1;
  offs[0] = 0;
  for (uint32_t i = 0; i < npos; i++) offs[i + 1] = offs[i] + (counts ? counts[i] : per_pos);
  int64_t total = (int64_t)offs[npos];
  (void)threads;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t g = 0; g < total; g++) {
    /* binary search for the position owning playout g */
    uint32_t lo = 0, hi = npos;
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) / 2;
      if (offs[mid] <= (uint64_t)g) lo = mid; else hi = mid;
    }
    uint32_t j = (uint32_t)((uint64_t)g - offs[lo]);
    or_trace t;
    int w = or_playout(pos[lo], seed, keys ? keys[lo] : lo, j, epoch, policy, traces ? &t : NULL);
    if (winners) winners[g] = (int8_t)w;
    if (traces) traces[g] = t;
  }
  free(offs);
  return 0;
}

/* src/driver.cpp:76-104, with a counter-based stream per position instead of the shared
 * engine: k uniform in [kmin, kmax] random plies, stopping early when the game is over. */
int or_gen_start_positions(uint32_t seed, uint32_t n, uint32_t kmin, uint32_t kmax,
                           int reject_terminal, or_pos *out) {
  if (kmax < kmin || kmax - kmin == 0xFFFFFFFFu) return -2;  /* kmax - kmin + 1 must fit 32 bits */
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint32_t pos = (uint32_t)i;
    board_state s;
    board_move list[LIST_CAP];
    int done = 0;
    for (uint32_t attempt = 0; attempt < 4096 && !done; attempt++) {
      or_pos start = or_start_position();
      unpack(&start, &s);
      uint32_t ctr_len[4] = {0, attempt, OR_TAG_START, 1}, key[2] = {seed, pos}, w[4];
      or_philox4x32_10(ctr_len, key, w);
      uint32_t k = kmin + w[0] % (kmax - kmin + 1);
      uint32_t words[4] = {0, 0, 0, 0};
      int n = gen(&s, list);
      for (uint32_t t = 0; t < k; t++) {
        if (game_over_with(&s, n)) break;
        if ((t & 3) == 0) {
          uint32_t ctr[4] = {t >> 2, attempt, OR_TAG_START, 0};
          or_philox4x32_10(ctr, key, words);
        }
        apply(&s, &list[words[t & 3] % (uint32_t)n]);
        n = gen(&s, list);
      }
      if (!reject_terminal || !game_over_with(&s, n)) done = 1;
    }
    if (!done) {
#pragma omp atomic write
      failed = 1;
    }
    out[pos] = pack(&s);
  }
  return failed ? -1 : 0;
}
