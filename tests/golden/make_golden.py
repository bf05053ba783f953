"""Writes the golden fixtures of tests/golden/ from the reference's own host code.

Run in the development container, after `make -C oracle ref` has compiled oracle/_ref from the
reference sources (src/state.cu, src/state.cpp, src/heuristic.cu, src/mcts.cpp) with asserts on:

    python tests/golden/make_golden.py

Every vector below is produced by the reference build (oracle/_ref/libref_assert.so):
  positions.npz   start positions of the benchmark configurations (genRandomStates-style,
                  src/driver.cpp:76-104, with the Philox stream of the oracle)
  movelists.npz   State::genMoves lists (src/state.cu:239-245) of ~4500 positions, flattened
  perft.json      move-tree leaf counts from the start position and from midgame positions
  playouts.npz    per-playout winner / ply count / move-list hash chain, random and heuristic
  scores.npz      scoreState / scoreMove / getWeight (src/heuristic.cu:20-49) samples
  full_configs.npz  every winner of the two full benchmark batches, 1024 positions x 1024
                  playouts (random on the k in [0, 30] openings, heuristic on the k in [20, 60]
                  midgame positions; seed 12345, epoch 0), two bits per winner (winner + 1),
                  and each position's summed ply count
  noisy.npz       the noisy heuristic (src/heuristicPlayout.cpp:12-48 with its N(0, 0.11)
                  noise): playout traces with the oracle's noise draws over the reference's own
                  move generation and weights; per-ply choice counts of the reference's weights
                  with its own std::normal_distribution noise; winner counts of the reference's
                  own HostHeuristicPlayoutDriver::runPlayouts
  MANIFEST.json   SHA-256 of each file

    python tests/golden/make_golden.py --full-only    # rewrites full_configs.npz only
    python tests/golden/make_golden.py --noisy-only   # rewrites noisy.npz only
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402

SEED_POS = 2016
SEED_PLAY = 12345


def pack2(w: np.ndarray) -> np.ndarray:
    """Winners (-1, 0, 1) as two-bit codes winner + 1, four per byte, first in the low bits."""
    c = (w.astype(np.int16) + 1).astype(np.uint8).reshape(-1, 4)
    return (c[:, 0] | (c[:, 1] << 2) | (c[:, 2] << 4) | (c[:, 3] << 6)).astype(np.uint8)


def full_configs(ref, open1024: np.ndarray, mid1024: np.ndarray) -> None:
    """The reference build's winners and per-position ply sums of the full benchmark batches."""
    out = {}
    for name, pos, policy in (("random_1024x1024", open1024, pyoracle.POLICY_RANDOM),
                              ("heuristic_1024x1024", mid1024, pyoracle.POLICY_HEURISTIC)):
        w, tr = ref.run_playouts(pos, per_pos=1024, seed=SEED_PLAY, epoch=0, policy=policy,
                                 threads=os.cpu_count() or 1)
        out[name + "_winners2"] = pack2(w)
        out[name + "_plies"] = tr["plies"].reshape(len(pos), 1024).sum(axis=1, dtype=np.uint64)
        print(f"{name}: {len(w)} playouts, {int(out[name + '_plies'].sum())} plies")
    np.savez_compressed(os.path.join(HERE, "full_configs.npz"), **out)


CHOICE_TRIALS = 20000
DRIVER_PLAYOUTS = 16384


def noisy_fixtures(ref, mid1024: np.ndarray, placed: np.ndarray) -> None:
    """noisy.npz: see the module docstring."""
    out = {}
    # playout traces: the reference's move lists and weights, the oracle's draws
    out["trace_pos"] = mid1024[:256]
    _, out["traces"] = ref.run_playouts(mid1024[:256], per_pos=16, seed=SEED_PLAY, epoch=0,
                                        policy=pyoracle.POLICY_HEURISTIC_NOISY)
    # per-ply choice counts: midgame positions whose moves do not all weigh the same (finite
    # weights), placements where two or more moves take the last opponent piece (weight +inf:
    # the first of them in reference order is played), and midgame positions whose moves all
    # weigh the same (the draws alone decide: a uniform choice)
    finite, infinite, single = [], [], []
    for p in ref.start_positions(SEED_POS + 1, 4000, 10, 70, True):
        ms = ref.gen_moves(p)
        ws = [ref.score_move_index(p, k)[2] for k in range(len(ms))]
        if len(ms) >= 2 and len(set(ws)) >= 2 and np.isfinite(max(ws)) and len(finite) < 64:
            finite.append(p)
        if len(ms) >= 3 and len(set(ws)) == 1 and np.isfinite(ws[0]) and len(single) < 16:
            single.append(p)
        if len(finite) == 64 and len(single) == 16:
            break
    for p in placed:
        if ref.is_game_over(p):
            continue
        ms = ref.gen_moves(p)
        ws = [ref.score_move_index(p, k)[2] for k in range(len(ms))]
        if sum(1 for w in ws if np.isinf(w)) >= 2:
            infinite.append(p)
        if len(infinite) == 16:
            break
    cpos = np.array(finite + infinite + single, dtype=pyoracle.POS_DTYPE)
    counts = np.zeros((len(cpos), 64), dtype=np.uint32)
    nmoves = np.zeros(len(cpos), dtype=np.uint32)
    for i, p in enumerate(cpos):
        c = ref.noisy_choice_counts(p, CHOICE_TRIALS, 1000 + i)
        counts[i, :len(c)] = c
        nmoves[i] = len(c)
    out.update(choice_pos=cpos, choice_counts=counts, choice_n=nmoves,
               choice_trials=np.uint32(CHOICE_TRIALS), choice_infinite_from=np.uint32(len(finite)),
               choice_single_from=np.uint32(len(finite) + len(infinite)))
    # the reference driver itself: winner counts (player 1, player 2, draw) per position
    dpos = mid1024[256:272]
    w = ref.host_heuristic_driver(dpos, DRIVER_PLAYOUTS).reshape(len(dpos), DRIVER_PLAYOUTS)
    out["driver_pos"] = dpos
    out["driver_counts"] = np.stack([(w == 0).sum(1), (w == 1).sum(1), (w == -1).sum(1)],
                                    axis=1).astype(np.uint32)
    np.savez_compressed(os.path.join(HERE, "noisy.npz"), **out)
    print(f"noisy: {len(out['traces'])} traces, {len(finite)} + {len(infinite)} + {len(single)} "
          f"choice positions, "
          f"driver {len(dpos)} x {DRIVER_PLAYOUTS}")


def write_manifest() -> None:
    manifest = {}
    for name in ("positions.npz", "movelists.npz", "perft.json", "playouts.npz", "scores.npz",
                 "full_configs.npz", "noisy.npz"):
        with open(os.path.join(HERE, name), "rb") as f:
            manifest[name] = hashlib.sha256(f.read()).hexdigest()
    manifest["source"] = "oracle/_ref/libref_assert.so (reference host code, asserts enabled)"
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1)


def main() -> None:
    ref = pyoracle.Reference(pyoracle.REF_ASSERT_SO)
    if "--full-only" in sys.argv:
        with np.load(os.path.join(HERE, "positions.npz")) as z:
            full_configs(ref, z["open1024"], z["mid1024"])
        write_manifest()
        return
    if "--noisy-only" in sys.argv:
        with np.load(os.path.join(HERE, "positions.npz")) as z:
            mid1024 = z["mid1024"]
        with np.load(os.path.join(HERE, "movelists.npz")) as z:
            positions = z["positions"]
        noisy_fixtures(ref, mid1024, positions)
        write_manifest()
        return

    # ---- start positions
    open1024 = ref.start_positions(SEED_POS, 1024, 0, 30, False)
    mid1024 = ref.start_positions(SEED_POS, 1024, 20, 60, True)
    np.savez_compressed(os.path.join(HERE, "positions.npz"), open1024=open1024, mid1024=mid1024)

    # ---- move lists: configuration starts plus positions sampled along seeded random games
    rng = np.random.Generator(np.random.Philox(7))
    pool = [open1024, mid1024]
    walk = []
    for start in np.concatenate([open1024[:400], mid1024[:400]]):
        p = start.copy()
        for _ in range(int(rng.integers(1, 120))):
            moves = ref.gen_moves(p)
            if len(moves) == 0 or p["counter"] >= 50:
                break
            walk.append(p.copy())
            p = ref.apply_index(p, int(rng.integers(0, len(moves))))
    walk = np.array(walk, dtype=pyoracle.POS_DTYPE)
    # keep every capture position and kings position, and a sample of the rest
    keep = []
    for i, p in enumerate(walk):
        ms = ref.gen_moves(p)
        if len(ms) and (ms[0]["jumps"] > 0 or p["kings"] or rng.random() < 0.05):
            keep.append(i)
    walk = walk[np.array(keep)]
    # random placements: any mix of men and kings of both sides, either side to move, draw
    # counters around the limit of 50 (src/state.cpp:16-23)
    placed = np.zeros(3000, dtype=pyoracle.POS_DTYPE)
    for i in range(len(placed)):
        occ = rng.random(32) < rng.uniform(0.1, 0.6)
        side = rng.random(32) < 0.5
        king = rng.random(32) < rng.uniform(0.0, 0.7)
        bits = lambda m: int(np.sum((1 << np.arange(32, dtype=np.uint64))[m]))
        placed[i]["p1"] = bits(occ & side)
        placed[i]["p2"] = bits(occ & ~side)
        placed[i]["kings"] = bits(occ & king)
        placed[i]["turn"] = int(rng.integers(0, 2))
        placed[i]["counter"] = int(rng.choice([0, 3, 48, 49, 50, 51]))
    # long king capture sequences (>= 4 jumps), where the landing-square cycle check of
    # src/state.cu:134-139 decides the list: one or two kings among many opponent men
    loops = []
    tries = 0
    while len(loops) < 300 and tries < 200000:
        tries += 1
        p = np.zeros(1, dtype=pyoracle.POS_DTYPE)[0]
        mover = int(rng.integers(0, 2))
        ks = rng.choice(32, size=int(rng.integers(1, 3)), replace=False)
        others = [q for q in rng.choice(32, size=int(rng.integers(5, 12)), replace=False) if q not in ks]
        own = sum(1 << int(q) for q in ks)
        opp = sum(1 << int(q) for q in others)
        p["p1" if mover == 0 else "p2"] = own
        p["p2" if mover == 0 else "p1"] = opp
        p["kings"] = own | (opp if rng.random() < 0.3 else 0)
        p["turn"] = mover
        ms = ref.gen_moves(p)
        if len(ms) and ms["jumps"].max() >= 4:
            loops.append(p)
    loops = np.array(loops, dtype=pyoracle.POS_DTYPE)
    positions = np.concatenate(pool + [walk, placed, loops])
    lists = [ref.gen_moves(p) for p in positions]
    counts = np.array([len(m) for m in lists], dtype=np.uint32)
    offsets = np.concatenate([[0], np.cumsum(counts, dtype=np.uint64)]).astype(np.uint64)
    moves = np.concatenate(lists) if len(lists) else np.zeros(0, dtype=pyoracle.MOVE_DTYPE)
    np.savez_compressed(os.path.join(HERE, "movelists.npz"), positions=positions, counts=counts,
                        offsets=offsets, moves=moves)

    # ---- perft
    perft = {"start": [ref.perft(pyoracle.start_position(), d) for d in range(1, 9)],
             "midgame": []}
    for p in mid1024[:8]:
        perft["midgame"].append({"position": [int(p["p1"]), int(p["p2"]), int(p["kings"]),
                                              int(p["turn"]), int(p["counter"])],
                                 "counts": [ref.perft(p, d) for d in range(1, 6)]})
    with open(os.path.join(HERE, "perft.json"), "w") as f:
        json.dump(perft, f, indent=1)

    # ---- playouts
    out = {}
    w, tr = ref.run_playouts(open1024[:256], per_pos=16, seed=SEED_PLAY, epoch=0,
                             policy=pyoracle.POLICY_RANDOM)
    out["random_open256x16"] = tr
    w, tr = ref.run_playouts(mid1024[:256], per_pos=16, seed=SEED_PLAY, epoch=0,
                             policy=pyoracle.POLICY_HEURISTIC)
    out["heuristic_mid256x16"] = tr
    counts_uneven = (np.arange(64, dtype=np.uint32) * 7) % 23
    w, tr = ref.run_playouts(mid1024[:64], counts=counts_uneven, seed=99, epoch=7,
                             policy=pyoracle.POLICY_RANDOM)
    out["random_mid64_uneven_e7"] = tr
    out["uneven_counts"] = counts_uneven
    np.savez_compressed(os.path.join(HERE, "playouts.npz"), **out)

    # ---- heuristic scores
    rows = []
    for p in positions[::7]:
        ms = ref.gen_moves(p)
        for k in range(len(ms)):
            s0, s1, wgt = ref.score_move_index(p, k)
            ss = ref.score_state(p)
            rows.append((p, k, s0, s1, wgt, ss[0], ss[1]))
    sc = np.zeros(len(rows), dtype=[("pos", pyoracle.POS_DTYPE), ("index", "<u4"),
                                    ("d0", "<i4"), ("d1", "<i4"), ("weight", "<f4"),
                                    ("s0", "<u4"), ("s1", "<u4")])
    for i, r in enumerate(rows):
        sc[i] = r
    np.savez_compressed(os.path.join(HERE, "scores.npz"), scores=sc)

    full_configs(ref, open1024, mid1024)
    noisy_fixtures(ref, mid1024, positions)
    write_manifest()
    print(f"positions {len(positions)}, moves {len(moves)}, max n {counts.max()}, "
          f"max jumps {moves['jumps'].max()}, king loops {len(loops)} ({tries} tries), "
          f"scores {len(sc)}")


if __name__ == "__main__":
    main()
