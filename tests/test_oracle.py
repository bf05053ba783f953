"""Pins the oracle (oracle/oracle.c) before it is trusted as the checker of the device path.

Checks, in order of strength:
  * published known answers: Random123's Philox4x32-10 vectors, English-draughts perft 1..8;
  * golden vectors written by the reference's own host code (tests/golden/, built from
    src/state.cu, src/state.cpp, src/heuristic.cu): move lists, perft, start positions,
    playout traces, heuristic scores;
  * closed-form positions (start list, promotion, double jump, draw counter, no-move loss);
  * mutation runs: the oracle rebuilt with one planted mistake must fail the golden checks.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def sq_mask(*squares):
    m = 0
    for s in squares:
        m |= 1 << s
    return m


def make_pos(p1=(), p2=(), kings=(), turn=0, counter=0):
    p = np.zeros(1, dtype=pyoracle.POS_DTYPE)[0]
    p["p1"] = sq_mask(*p1)
    p["p2"] = sq_mask(*p2)
    p["kings"] = sq_mask(*kings)
    p["turn"] = turn
    p["counter"] = counter
    return p


# ---------------------------------------------------------------------------------------------
# known answers
# ---------------------------------------------------------------------------------------------

# Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key, output)
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, expect):
    assert list(oracle.philox(ctr, key)) == expect


def test_stream_word_layout(oracle):
    # word (t & 3) of counter (t >> 2, playout, tag, epoch) under key (seed, position)
    for t in (0, 1, 2, 3, 4, 7, 100):
        block = oracle.philox([t >> 2, 5, 0, 9], [12345, 3])
        assert oracle.stream_word(12345, 3, t, 5, 0, 9) == block[t & 3]


# Published perft of 8x8 English draughts from the initial position (no kings can arise, so the
# reference's king-capture rule variants cannot matter at these depths).
PERFT_PUBLISHED = [7, 49, 302, 1469, 7361, 36768, 179740, 845931]


def test_perft_start_published(oracle):
    start = pyoracle.start_position()
    assert [oracle.perft(start, d) for d in range(1, 9)] == PERFT_PUBLISHED


def test_perft_golden(oracle, golden):
    assert golden["perft"]["start"] == PERFT_PUBLISHED
    for entry in golden["perft"]["midgame"]:
        p1, p2, kings, turn, counter = entry["position"]
        p = make_pos(turn=turn, counter=counter)
        p["p1"], p["p2"], p["kings"] = p1, p2, kings
        assert [oracle.perft(p, d) for d in range(1, 6)] == entry["counts"]


# ---------------------------------------------------------------------------------------------
# golden vectors from the reference build
# ---------------------------------------------------------------------------------------------

def compare_movelists(lib, golden, limit=None) -> int:
    ml = golden["movelists"]
    pos, counts, offs, moves = ml["positions"], ml["counts"], ml["offsets"], ml["moves"]
    n = len(pos) if limit is None else min(limit, len(pos))
    mine, mine_counts = lib.gen_moves_many(pos[:n], cap=64)
    bad = 0
    for i in range(n):
        c = int(counts[i])
        if int(mine_counts[i]) != c:
            bad += 1
            continue
        ref = moves[int(offs[i]):int(offs[i]) + c]
        if c and mine[i, :c].tobytes() != ref.tobytes():
            bad += 1
    return bad


def test_movelists_golden(oracle, golden):
    assert compare_movelists(oracle, golden) == 0
    assert len(golden["movelists"]["positions"]) > 20000


def test_start_positions_golden(oracle, golden):
    open1024 = oracle.start_positions(2016, 1024, 0, 30, False)
    mid1024 = oracle.start_positions(2016, 1024, 20, 60, True)
    assert open1024.tobytes() == golden["positions"]["open1024"].tobytes()
    assert mid1024.tobytes() == golden["positions"]["mid1024"].tobytes()
    # the midgame set rejects finished positions
    assert not any(oracle.is_game_over(p) for p in mid1024)
    # kmax < kmin and a 2^32-wide ply range are refused (the range would wrap to zero)
    for kmin, kmax in ((5, 4), (0, 2**32 - 1)):
        with pytest.raises(RuntimeError):
            oracle.start_positions(3, 4, kmin, kmax)


def playout_sets(golden):
    g, pl = golden["positions"], golden["playouts"]
    return [
        (g["open1024"][:256], dict(per_pos=16, seed=12345, epoch=0, policy=0), pl["random_open256x16"]),
        (g["mid1024"][:256], dict(per_pos=16, seed=12345, epoch=0, policy=1), pl["heuristic_mid256x16"]),
        (g["mid1024"][:64], dict(counts=pl["uneven_counts"], seed=99, epoch=7, policy=0),
         pl["random_mid64_uneven_e7"]),
    ]


def compare_traces(a, b) -> int:
    return int(np.sum((a["winner"] != b["winner"]) | (a["plies"] != b["plies"]) |
                      (a["hash"] != b["hash"]) | (a["max_n"] != b["max_n"])))


def test_playouts_golden(oracle, golden):
    for pos, kw, expect in playout_sets(golden):
        _, tr = oracle.run_playouts(pos, **kw)
        assert compare_traces(tr, expect) == 0


def test_scores_golden(oracle, golden):
    sc = golden["scores"]
    for row in sc[::5]:
        p = row["pos"]
        moves = oracle.gen_moves(p)
        m = moves[int(row["index"])]
        assert oracle.score_state(p) == (int(row["s0"]), int(row["s1"]))
        assert oracle.score_move(p, m) == (int(row["d0"]), int(row["d1"]))


def test_exact_weight_order_matches_reference_floats(golden):
    """The oracle ranks (S_m + d_m) / (S_o + d_o) exactly; the reference compares floats
    (src/heuristic.cu:44-49).  Within each position both orders must agree, ties included."""
    sc = golden["scores"]
    keys = sc["pos"].tobytes()
    start = 0
    n = len(sc)
    checked = 0
    while start < n:
        end = start
        while end < n and sc["pos"][end].tobytes() == sc["pos"][start].tobytes():
            end += 1
        rows = sc[start:end]
        turn = int(rows["pos"]["turn"][0])
        S = [int(rows["s0"][0]), int(rows["s1"][0])]
        d = np.stack([rows["d0"], rows["d1"]], axis=1).astype(np.int64)
        num = S[turn] + d[:, turn]
        den = S[1 - turn] + d[:, 1 - turn]
        w = rows["weight"]
        for i in range(len(rows)):
            for j in range(len(rows)):
                exact = np.sign(num[i] * den[j] - num[j] * den[i])
                flt = np.sign(float(w[i]) - float(w[j])) if not (np.isinf(w[i]) and np.isinf(w[j])) else 0
                assert exact == flt
                checked += 1
        start = end
    del keys
    assert checked > 1000


# ---------------------------------------------------------------------------------------------
# closed-form positions
# ---------------------------------------------------------------------------------------------

def moves_as_tuples(ms):
    return [(int(m["from"]), int(m["to"]), int(m["jumps"]), int(m["promoted"]),
             tuple(int(x) for x in m["inter"][:m["jumps"]]),
             tuple(int(x) for x in m["removed"][:m["jumps"]])) for m in ms]


def test_start_list_closed_form(oracle):
    # rows 2 -> 3, each man's (+1,+1) step before its (+1,-1) step (src/state.cu:255-281)
    got = [(a, b) for a, b, *_ in moves_as_tuples(oracle.gen_moves(pyoracle.start_position()))]
    assert got == [(8, 13), (8, 12), (9, 14), (9, 13), (10, 15), (10, 14), (11, 15)]


def test_double_jump_closed_form(oracle):
    # man on b8... square 1 = (0,3); opponents on (1,2)=5 and (3,2)=13: 1 -> 8 -> 17
    p = make_pos(p1=[1], p2=[5, 13, 31], turn=0)
    assert moves_as_tuples(oracle.gen_moves(p)) == [(1, 17, 2, 0, (8, 17), (5, 13))]
    q = oracle.apply(p, oracle.gen_moves(p)[0])
    assert int(q["p1"]) == 1 << 17 and int(q["p2"]) == 1 << 31 and int(q["counter"]) == 0


def test_promotion_closed_form(oracle):
    # player-1 man on row 6 square 24 = (6,1) steps to (7,2)=29 or (7,0)=28, both promoted
    p = make_pos(p1=[24], p2=[3], turn=0, counter=7)
    ms = moves_as_tuples(oracle.gen_moves(p))
    assert ms == [(24, 29, 0, 1, (), ()), (24, 28, 0, 1, (), ())]
    q = oracle.apply(p, oracle.gen_moves(p)[0])
    assert int(q["kings"]) == 1 << 29 and int(q["counter"]) == 8 and int(q["turn"]) == 1


def test_draw_counter_closed_form(oracle):
    p = make_pos(p1=[9], p2=[22], turn=0, counter=49)
    assert not oracle.is_game_over(p)
    p["counter"] = 50
    assert oracle.is_game_over(p) and oracle.winner(p) == -1


def test_no_moves_loses(oracle):
    p = make_pos(p1=[], p2=[22], turn=0)
    assert oracle.is_game_over(p) and oracle.winner(p) == 1
    # blocked: player-2 man on 28 = (7,0); its only step (6,1) = 24 is taken and the jump
    # square (5,2) = 21 is taken too
    p = make_pos(p1=[24, 21], p2=[28], turn=1)
    assert len(oracle.gen_moves(p)) == 0 and oracle.winner(p) == 0


def test_king_sequences_never_revisit(oracle, golden):
    ml = golden["movelists"]
    for m in ml["moves"]:
        j = int(m["jumps"])
        if j > 1:
            lands = list(m["inter"][:j])
            assert len(set(lands)) == j and int(m["from"]) not in lands


# ---------------------------------------------------------------------------------------------
# mutation runs: each planted mistake must be caught by a golden check
# ---------------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")
    libs = {}
    src = os.path.join(ROOT, "oracle", "oracle.c")
    for k in range(1, 7):
        out = str(d / f"liboracle_m{k}.so")
        subprocess.run(["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared",
                        f"-DOR_MUTATION={k}", src, "-o", out], check=True)
        libs[k] = pyoracle.Oracle(out)
    return libs


@pytest.mark.parametrize("k", range(1, 7))
def test_mutation_is_detected(mutants, golden, k):
    lib = mutants[k]
    caught = compare_movelists(lib, golden) > 0
    for pos, kw, expect in playout_sets(golden):
        _, tr = lib.run_playouts(pos, **kw)
        caught |= compare_traces(tr, expect) > 0
    if k == 6:
        caught |= list(lib.philox(*PHILOX_KAT[0][:2])) != PHILOX_KAT[0][2]
    assert caught, f"planted mistake {k} went unnoticed"


# ---------------------------------------------------------------------------------------------
# direct comparison with the reference build (development container only)
# ---------------------------------------------------------------------------------------------

@pytest.mark.skipif(not pyoracle.reference_available(), reason="oracle/_ref not built here")
def test_oracle_matches_reference_build_64_positions(oracle):
    ref = pyoracle.Reference()
    pos = oracle.start_positions(2016, 64, 0, 30, False)
    assert pos.tobytes() == ref.start_positions(2016, 64, 0, 30, False).tobytes()
    for policy in (0, 1, 2):
        _, a = oracle.run_playouts(pos, per_pos=64, seed=4242, epoch=3, policy=policy)
        _, b = ref.run_playouts(pos, per_pos=64, seed=4242, epoch=3, policy=policy)
        assert compare_traces(a, b) == 0


@pytest.mark.parametrize("name", ["random_1024x1024", "heuristic_1024x1024"])
def test_oracle_full_configs_match_reference_golden(oracle, golden, name):
    """The restatement on the full benchmark batches (1024 x 1024 playouts, seed 12345, epoch 0)
    against every winner and per-position ply sum the reference build wrote."""
    policy = 0 if name.startswith("random") else 1
    pos = golden["positions"]["open1024" if policy == 0 else "mid1024"]
    w, tr = oracle.run_playouts(pos, per_pos=1024, seed=12345, epoch=0, policy=policy)
    expect = golden["full"][name]
    assert np.array_equal(w, expect["winners"])
    plies = tr["plies"].reshape(1024, 1024).sum(axis=1, dtype=np.uint64)
    assert np.array_equal(plies, expect["plies"])


def test_oracle_stream_keys(oracle, golden):
    """Keyed playouts: keys equal to the position indices are the unkeyed playouts; a keyed
    position plays the same games alone as inside a batch."""
    pos = golden["positions"]["open1024"][:64]
    counts = (np.arange(64, dtype=np.uint32) * 5) % 17
    plain, _ = oracle.run_playouts(pos, counts=counts, seed=3, epoch=2, traces=False)
    same, _ = oracle.run_playouts(pos, counts=counts, seed=3, epoch=2, traces=False,
                                  keys=np.arange(64, dtype=np.uint32))
    assert np.array_equal(plain, same)
    keys = (np.arange(64, dtype=np.uint32) * 2654435761) & 0xFFFFFFFF
    allk, _ = oracle.run_playouts(pos, counts=counts, seed=3, epoch=2, traces=False, keys=keys)
    offs = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])
    for i in (5, 40):
        one, _ = oracle.run_playouts(pos[i:i + 1], counts=counts[i:i + 1], seed=3, epoch=2,
                                     traces=False, keys=keys[i:i + 1])
        assert np.array_equal(one, allk[offs[i]:offs[i + 1]])
    # a different key is a different stream
    other, _ = oracle.run_playouts(pos, counts=counts, seed=3, epoch=2, traces=False, keys=keys + 1)
    assert not np.array_equal(other, allk)
