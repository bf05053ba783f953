"""The launch-shape table (include/gpuai/mi_playout.h, mi_autotune / mi_autotune_sizes): each
batch-size bucket holds the shape its tuning run measured fastest, mi_run_playouts and graph
replays launch each batch with its bucket's shape, and winners do not depend on the shape.

The reference picks a playout driver by batch size with a runtime model
(OptimalPlayoutDriver, src/playout.cpp:80-166); here the choice is a measured table.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from gpu_ai_amd.engine import POLICY_RANDOM, REFILL_QUEUE, Engine

pytestmark = pytest.mark.gpu

SIZES = [(64, 16), (256, 64), (1024, 256)]  # (positions, playouts each): buckets 10, 14, 18


def announce(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def bucket(total: int) -> int:
    return total.bit_length() - 1


@pytest.fixture(scope="module")
def tuned():
    eng = Engine(0)
    pos = eng.start_positions(2016, 1024, 0, 30)
    rows = {}
    for n, per in SIZES:
        announce(f"mi_autotune on {n} positions x {per} playouts")
        _, rows[n * per] = eng.autotune(pos[:n], per_pos=per, reps=3)
    yield eng, pos, rows
    eng.close()


def test_each_bucket_holds_its_fastest_shape(tuned):
    eng, _, rows = tuned
    table = {(e["bucket"], e["refill"]): e for e in eng.tuning_table() if e["policy"] == POLICY_RANDOM}
    for total, cand in rows.items():
        for refill in (0, 1):
            mine = [r for r in cand if r["refill"] == refill]
            fastest = min(mine, key=lambda r: r["kernel_ms"])
            e = table[(bucket(total), refill)]
            assert e["playouts"] == total
            assert abs(e["kernel_ms"] - fastest["kernel_ms"]) < 1e-6
            assert (e["block_size"], e["blocks_per_cu"], e["defer_min"]) == \
                (fastest["block_size"], fastest["blocks_per_cu"], fastest["defer_min"])


def test_launches_take_their_bucket_shape(tuned):
    eng, pos, _ = tuned
    table = {(e["bucket"], e["refill"]): e for e in eng.tuning_table() if e["policy"] == POLICY_RANDOM}
    with Engine(0) as fixed:
        fixed.set_tuning(256, 4, REFILL_QUEUE, 8)  # one explicit shape for every size
        for n, per in SIZES:
            total = n * per
            announce(f"mi_run_playouts {n}x{per} with the table, and with a fixed shape")
            r = eng.run_playouts(pos[:n], per_pos=per, seed=77, epoch=3)
            best = min((table[(bucket(total), rf)] for rf in (0, 1)), key=lambda e: e["kernel_ms"])
            assert r.stats["block_size"] == best["block_size"]
            assert r.stats["refill"] == best["refill"]
            ref = fixed.run_playouts(pos[:n], per_pos=per, seed=77, epoch=3)
            assert ref.stats["block_size"] == 256
            assert np.array_equal(r.winners, ref.winners)


def test_graph_replays_take_their_bucket_shape(tuned):
    eng, pos, _ = tuned
    table = {(e["bucket"], e["refill"]): e for e in eng.tuning_table() if e["policy"] == POLICY_RANDOM}
    big = SIZES[-1][0] * SIZES[-1][1]
    g = eng.graph(1024, big, POLICY_RANDOM)
    try:
        refill = None
        for n, per in reversed(SIZES):
            total = n * per
            announce(f"graph replay {n}x{per}")
            counts = np.full(n, per, dtype=np.uint32)
            r = g.run(pos[:n], counts, seed=5, epoch=1)
            refill = r.stats["refill"] if refill is None else refill
            assert r.stats["refill"] == refill  # a graph keeps its refill policy
            e = table[(bucket(total), refill)]
            assert r.stats["block_size"] == e["block_size"]
            direct = eng.run_playouts(pos[:n], counts=counts, seed=5, epoch=1)
            assert np.array_equal(r.winners, direct.winners)
    finally:
        g.close()


def test_autotune_sizes_and_set_tuning(tuned):
    _, pos, _ = tuned
    with Engine(0) as eng:
        announce("mi_autotune_sizes 3000, 50000")
        best = eng.autotune_sizes(pos, [3000, 50000], reps=1)
        assert [b["playouts"] for b in best] == [3000 // 1024 * 1024, 50000 // 1024 * 1024]
        table = eng.tuning_table()
        assert {(e["bucket"], e["refill"]) for e in table} == {(11, 0), (11, 1), (15, 0), (15, 1)}
        for b in best:
            same = [e for e in table if e["bucket"] == b["bucket"]]
            assert b["kernel_ms"] == min(e["kernel_ms"] for e in same)
        r = eng.run_playouts(pos[:1024], per_pos=2, seed=1)  # bucket 11
        assert r.stats["block_size"] == best[0]["block_size"]
        eng.set_tuning(128, 2, REFILL_QUEUE, 16)
        assert eng.tuning_table() == []
        assert eng.run_playouts(pos[:1024], per_pos=2, seed=1).stats["block_size"] == 128


def test_table_keeps_policies_apart(tuned):
    """Each policy has its own rows of the table: tuning the noisy heuristic at a size leaves the
    random policy's entry of that bucket as it was, and each launch takes its own policy's shape
    with results equal to those of a fixed shape."""
    from gpu_ai_amd.engine import POLICY_HEURISTIC_NOISY
    _, pos, _ = tuned
    with Engine(0) as eng:
        announce("mi_autotune_sizes 16384, random then noisy heuristic")
        rnd = eng.autotune_sizes(pos, [16384], policy=POLICY_RANDOM, reps=1)[0]
        before = {(e["policy"], e["bucket"], e["refill"]): e for e in eng.tuning_table()}
        noisy = eng.autotune_sizes(pos, [16384], policy=POLICY_HEURISTIC_NOISY, reps=1)[0]
        table = {(e["policy"], e["bucket"], e["refill"]): e for e in eng.tuning_table()}
        assert {p for p, _, _ in table} == {POLICY_RANDOM, POLICY_HEURISTIC_NOISY}
        for key, e in before.items():
            assert table[key] == e  # the random rows are untouched
        assert rnd["policy"] == POLICY_RANDOM and noisy["policy"] == POLICY_HEURISTIC_NOISY
        sel = pos[:1024]  # 1024 x 16 playouts: the tuned bucket
        for policy, best in ((POLICY_RANDOM, rnd), (POLICY_HEURISTIC_NOISY, noisy)):
            r = eng.run_playouts(sel, per_pos=16, seed=3, epoch=1, policy=policy)
            assert (r.stats["block_size"], r.stats["refill"]) == \
                (best["block_size"], best["refill"]), policy
            with Engine(0) as fixed:
                fixed.set_tuning(256, 4, REFILL_QUEUE, 8)
                ref = fixed.run_playouts(sel, per_pos=16, seed=3, epoch=1, policy=policy)
            assert np.array_equal(r.winners, ref.winners), policy
