/* oracle_sanitize_check.c -- drives every entry point of the oracle's C restatement
 * (oracle/oracle.c) over seeded inputs and prints one digest of all the results.
 * tests/test_sanitizers.py builds it twice, plainly and under AddressSanitizer +
 * UndefinedBehaviorSanitizer, and requires the same digest from both builds. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint32_t digest = 0;

static void mix(uint32_t v) { digest = or_fmix32(digest ^ v) + 0x9E3779B9u; }

static void mix_traces(const or_trace *t, const int8_t *w, size_t n) {
  for (size_t i = 0; i < n; i++) {
    mix((uint32_t)(uint8_t)w[i]);
    mix((uint32_t)(uint8_t)t[i].winner | ((uint32_t)t[i].max_n << 8));
    mix(t[i].plies);
    mix(t[i].hash);
    mix(t[i].captures);
  }
}

int main(void) {
  enum { NPOS = 192, PER = 8 };
  static or_pos open[NPOS], mid[NPOS];
  if (or_gen_start_positions(2016, NPOS, 0, 30, 0, open) != 0) return 2;
  if (or_gen_start_positions(2016, NPOS, 20, 60, 1, mid) != 0) return 2;
  if (or_gen_start_positions(3, 4, 5, 4, 0, open) != -2) return 2; /* kmax < kmin refused */
  for (int i = 0; i < NPOS; i++) {
    mix(open[i].p1), mix(open[i].p2), mix(open[i].kings), mix(open[i].turn), mix(open[i].counter);
    mix(mid[i].p1), mix(mid[i].p2), mix(mid[i].kings), mix(mid[i].turn), mix(mid[i].counter);
  }

  /* move lists, every move applied, scores and digests */
  static or_move list[OR_MAX_MOVES];
  for (int i = 0; i < NPOS; i++) {
    const or_pos *src = (i & 1) ? &mid[i] : &open[i];
    const int n = or_gen_moves(src, list, OR_MAX_MOVES);
    if (n < 0) return 3;
    mix((uint32_t)n);
    mix(or_list_digest(list, n));
    for (int m = 0; m < n; m++) {
      or_pos q = *src;
      int32_t ds[2];
      or_score_move(src, &list[m], ds);
      mix((uint32_t)ds[0]), mix((uint32_t)ds[1]);
      if (or_apply(&q, &list[m]) != 0) return 3;
      mix(q.p1 ^ q.p2 ^ q.kings), mix((uint32_t)or_is_game_over(&q));
      if (or_is_game_over(&q)) mix((uint32_t)or_winner(&q));
    }
    uint32_t sc[2];
    or_score_state(src, sc);
    mix(sc[0]), mix(sc[1]);
  }

  /* playouts of every policy, uniform and uneven counts, with and without stream keys */
  static int8_t w[NPOS * PER * 2];
  static or_trace tr[NPOS * PER * 2];
  static uint32_t counts[NPOS], keys[NPOS];
  uint32_t total = 0;
  for (int i = 0; i < NPOS; i++) {
    counts[i] = (uint32_t)((i * 37) % 15);
    keys[i] = or_fmix32((uint32_t)i * 2654435761u);
    total += counts[i];
  }
  const int policies[] = {OR_POLICY_RANDOM, OR_POLICY_HEURISTIC, OR_POLICY_HEURISTIC_NOISY,
                          OR_POLICY_RANDOM_LIBC};
  for (int k = 0; k < 4; k++) {
    const int pol = policies[k];
    if (or_run_playouts(pol == OR_POLICY_RANDOM ? open : mid, NPOS, NULL, PER, 12345, 1, pol, w,
                        tr, 1) != 0)
      return 4;
    if (pol != OR_POLICY_RANDOM_LIBC) mix_traces(tr, w, (size_t)NPOS * PER);
    if (pol == OR_POLICY_RANDOM_LIBC) continue;
    if (or_run_playouts_keyed(mid, NPOS, counts, 0, keys, 99, 7, pol, w, tr, 1) != 0) return 4;
    mix_traces(tr, w, total);
  }

  /* one ply's choice frequencies and the noise draws */
  static uint32_t freq[OR_MAX_MOVES];
  for (int i = 0; i < 24; i++)
    for (int pol = 0; pol < 3; pol++) {
      const int n = or_choice_counts(&mid[i], pol, 5, 2000, freq, OR_MAX_MOVES);
      if (n < 0) return 5;
      for (int m = 0; m < n; m++) mix(freq[m]);
    }
  static uint32_t a[64], b[64];
  static float z[128];
  for (int i = 0; i < 64; i++) {
    a[i] = i < 4 ? (uint32_t)(i ? 0xFFFFFFFFu : 0u) : or_fmix32((uint32_t)i);
    b[i] = i < 4 ? (uint32_t)(i & 1 ? 0xFFFFFFFFu : 0u) : or_fmix32((uint32_t)i + 1000u);
  }
  or_normal_pairs(a, b, 64, z);
  for (int i = 0; i < 128; i++) {
    uint32_t bits;
    memcpy(&bits, &z[i], sizeof bits);
    mix(bits);
  }
  for (uint32_t m = 0; m < 16; m++) {
    const float v = or_heuristic_noise(or_fmix32(m), m);
    uint32_t bits;
    memcpy(&bits, &v, sizeof bits);
    mix(bits);
  }
  printf("digest %08x\n", digest);
  return 0;
}
