"""Parity of the gfx950 playout path with the oracle and with the reference's golden vectors.

Every test calls the device through the C ABI (gpu_ai_amd.engine -> libgpuai_mi355x.so).  The
bar is bit-exact: move lists, winners, ply counts and the per-playout hash chain over every
generated move list and every chosen index.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from gpu_ai_amd.engine import (POLICY_HEURISTIC, POLICY_HEURISTIC_NOISY, POLICY_RANDOM,
                               REFILL_QUEUE, REFILL_STATIC)

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def trace_mismatches(res, tr) -> int:
    return int(np.sum((res.winners != tr["winner"]) | (res.plies != tr["plies"]) |
                      (res.hashes != tr["hash"])))


def test_divmod_exact(engine):
    announce("mi_check_divmod: n in [1, 1100] and large n, boundary and random r")
    rng = np.random.default_rng(1)
    ns, rs = [], []
    large = [2**17 - 1, 2**17, 2**18 - 1, 2**18, 2**20 + 7, 2**31, 2**32 - 1]
    for n in list(range(1, 1101)) + large:
        base = [0, 1, n - 1, n, n + 1, 2**31 - 1, 2**31, 2**32 - 1, 2**32 - n, 2**32 - n - 1]
        q = (2**32 - 1) // n
        base += [q * n - 1, q * n, (q // 2) * n, (q // 2) * n - 1]
        # exact multiples and the values just below them (r/n an integer, or 1/n short of one)
        k = rng.integers(1, q + 1, 32, dtype=np.uint64)
        base += [int(x) * n for x in k] + [int(x) * n - 1 for x in k]
        r = np.array([x % 2**32 for x in base] + list(rng.integers(0, 2**32, 64, dtype=np.uint64)),
                     dtype=np.uint64)
        rs.append(r)
        ns.append(np.full(len(r), n, dtype=np.uint64))
    r = np.concatenate(rs).astype(np.uint32)
    n = np.concatenate(ns).astype(np.uint32)
    q, rem = engine.check_divmod(r, n)
    assert np.array_equal(q, r // n)
    assert np.array_equal(rem, r % n)


def test_gen_moves_golden(engine, golden):
    ml = golden["movelists"]
    pos, counts, offs, moves = ml["positions"], ml["counts"], ml["offsets"], ml["moves"]
    announce(f"mi_gen_moves on {len(pos)} golden positions")
    got, got_counts = engine.gen_moves(pos, cap=32)
    assert np.array_equal(got_counts, counts)
    flat = np.concatenate([got[i, :counts[i]] for i in range(len(pos))])
    assert flat.tobytes() == moves.tobytes()


def test_unknown_policy_is_rejected(engine, golden):
    """Policy numbers past MI_POLICY_HEURISTIC_NOISY are refused by every entry point that takes
    one, before anything runs, and the context stays usable."""
    from gpu_ai_amd.engine import MI_ERR_INVALID_ARGUMENT, MiError
    pos = golden["positions"]["open1024"][:8]
    announce("unknown policy 3 at run, traced, autotune and graph creation")
    for call in (lambda: engine.run_playouts(pos, per_pos=4, policy=3),
                 lambda: engine.run_traced(pos, per_pos=4, policy=3),
                 lambda: engine.autotune(pos, per_pos=4, policy=3, reps=1),
                 lambda: engine.graph(8, 64, policy=3)):
        with pytest.raises(MiError) as e:
            call()
        assert e.value.code == MI_ERR_INVALID_ARGUMENT
    ok = engine.run_playouts(pos, per_pos=4, policy=POLICY_HEURISTIC_NOISY)
    assert ok.winners.size == 32


def test_gen_moves_capacity_error(engine, golden):
    from gpu_ai_amd.engine import MI_ERR_CAPACITY, MiError
    pos = golden["positions"]["open1024"][:64]
    with pytest.raises(MiError) as e:
        engine.gen_moves(pos, cap=1)
    assert e.value.code == MI_ERR_CAPACITY


def golden_sets(golden):
    g, pl = golden["positions"], golden["playouts"]
    return [
        ("random open1024[:256]x16", g["open1024"][:256], dict(per_pos=16, seed=12345, epoch=0, policy=0),
         pl["random_open256x16"]),
        ("heuristic mid1024[:256]x16", g["mid1024"][:256], dict(per_pos=16, seed=12345, epoch=0, policy=1),
         pl["heuristic_mid256x16"]),
        ("random mid1024[:64] uneven e7", g["mid1024"][:64],
         dict(counts=pl["uneven_counts"], seed=99, epoch=7, policy=0), pl["random_mid64_uneven_e7"]),
    ]


def test_traced_playouts_match_reference_golden(engine, golden):
    for name, pos, kw, expect in golden_sets(golden):
        announce(f"traced {name}")
        res = engine.run_traced(pos, **kw)
        assert trace_mismatches(res, expect) == 0, name
        assert res.stats["plies"] == int(expect["plies"].sum(dtype=np.uint64))


@pytest.mark.parametrize("policy", [POLICY_RANDOM, POLICY_HEURISTIC, POLICY_HEURISTIC_NOISY])
def test_traced_full_positions_vs_oracle(engine, oracle, golden, policy):
    pos = golden["positions"]["open1024" if policy == POLICY_RANDOM else "mid1024"]
    announce(f"traced 1024 positions x 64, policy {policy}")
    res = engine.run_traced(pos, per_pos=64, seed=2024, epoch=1, policy=policy)
    _, tr = oracle.run_playouts(pos, per_pos=64, seed=2024, epoch=1, policy=policy)
    assert trace_mismatches(res, tr) == 0
    assert res.stats["max_moves"] == int(tr["max_n"].max())


def test_random_1024x1024_full_bitexact(engine, oracle, golden):
    pos = golden["positions"]["open1024"]
    announce("random-1024x1024: 1024 positions x 1024 random playouts")
    res = engine.run_playouts(pos, per_pos=1024, seed=12345, epoch=0, policy=POLICY_RANDOM)
    w, tr = oracle.run_playouts(pos, per_pos=1024, seed=12345, epoch=0, policy=POLICY_RANDOM)
    assert np.array_equal(res.winners, w)
    assert res.stats["plies"] == int(tr["plies"].sum(dtype=np.uint64))
    # every winner the reference build wrote for this batch
    expect = golden["full"]["random_1024x1024"]
    assert np.array_equal(res.winners, expect["winners"])
    assert res.stats["plies"] == int(expect["plies"].sum(dtype=np.uint64))


def test_heuristic_1024x1024_full_bitexact(engine, oracle, golden):
    pos = golden["positions"]["mid1024"]
    announce("heuristic-1024x1024: 1024 midgame positions x 1024 heuristic playouts")
    res = engine.run_playouts(pos, per_pos=1024, seed=12345, epoch=0, policy=POLICY_HEURISTIC)
    w, tr = oracle.run_playouts(pos, per_pos=1024, seed=12345, epoch=0, policy=POLICY_HEURISTIC)
    assert np.array_equal(res.winners, w)
    assert res.stats["plies"] == int(tr["plies"].sum(dtype=np.uint64))
    expect = golden["full"]["heuristic_1024x1024"]
    assert np.array_equal(res.winners, expect["winners"])
    assert res.stats["plies"] == int(expect["plies"].sum(dtype=np.uint64))


def test_noisy_traces_match_reference_golden(engine, golden):
    """The noisy heuristic's traces written by the reference's own move generation and weights
    with the same draws (256 midgame positions x 16)."""
    g = golden["noisy"]
    announce("noisy heuristic traced, 256 x 16, against the reference's traces")
    res = engine.run_traced(g["trace_pos"], per_pos=16, seed=12345, epoch=0,
                            policy=POLICY_HEURISTIC_NOISY)
    assert trace_mismatches(res, g["traces"]) == 0
    assert res.stats["max_moves"] == int(g["traces"]["max_n"].max())


def test_noisy_1024x1024_full_bitexact(engine, oracle, golden):
    pos = golden["positions"]["mid1024"]
    announce("noisy-1024x1024: 1024 midgame positions x 1024 noisy heuristic playouts")
    res = engine.run_playouts(pos, per_pos=1024, seed=12345, epoch=0,
                              policy=POLICY_HEURISTIC_NOISY)
    w, tr = oracle.run_playouts(pos, per_pos=1024, seed=12345, epoch=0,
                               policy=POLICY_HEURISTIC_NOISY)
    assert np.array_equal(res.winners, w)
    assert res.stats["plies"] == int(tr["plies"].sum(dtype=np.uint64))


def test_noisy_outcomes_match_reference_driver(engine, golden):
    """Winner counts of the noisy policy on the device against those of the reference's own
    HostHeuristicPlayoutDriver (16 midgame positions x 16384 playouts, its own noise), by a
    pooled two-sample chi-square test; random playouts from the same positions fail it."""
    from scipy import stats
    g = golden["noisy"]
    per = 16384

    def pooled_p(policy):
        w = engine.run_playouts(g["driver_pos"], per_pos=per, seed=321, epoch=0,
                                policy=policy).winners.reshape(-1, per)
        chi, dof = 0.0, 0
        for ref, row in zip(g["driver_counts"], w):
            mine = np.array([(row == 0).sum(), (row == 1).sum(), (row == -1).sum()])
            t = np.stack([ref, mine]).astype(np.float64)
            t = t[:, t.sum(axis=0) > 0]
            if t.shape[1] < 2:
                continue
            r = stats.chi2_contingency(t, correction=False)
            chi += r.statistic
            dof += r.dof
        return float(stats.chi2.sf(chi, dof))

    announce("noisy heuristic 16 x 16384 against the reference driver's winner counts")
    assert pooled_p(POLICY_HEURISTIC_NOISY) > 1e-3
    assert pooled_p(POLICY_RANDOM) < 1e-12


def test_random_64x4096_bitexact(engine, oracle, golden):
    pos = golden["positions"]["open1024"][:64]
    announce("random-64x4096: 64 positions x 4096 random playouts")
    res = engine.run_playouts(pos, per_pos=4096, seed=12345, epoch=0, policy=POLICY_RANDOM)
    w, _ = oracle.run_playouts(pos, per_pos=4096, seed=12345, epoch=0, policy=POLICY_RANDOM,
                               traces=False)
    assert np.array_equal(res.winners, w)


@pytest.mark.parametrize("block,bpc,refill,defer", [(64, 1, REFILL_STATIC, 1), (64, 16, REFILL_QUEUE, 16),
                                                    (256, 8, REFILL_STATIC, 8), (1024, 2, REFILL_QUEUE, 1),
                                                    (512, 4, REFILL_QUEUE, 64)])
@pytest.mark.parametrize("policy", [POLICY_RANDOM, POLICY_HEURISTIC, POLICY_HEURISTIC_NOISY])
def test_launch_shape_does_not_change_results(engine, golden, block, bpc, refill, defer, policy):
    pos = golden["positions"]["mid1024"][:256]
    announce(f"tuning block {block} blocks/CU {bpc} refill {refill} defer {defer} policy {policy}")
    base = engine.run_traced(pos, per_pos=256, seed=7, epoch=2, policy=policy)
    fast = engine.run_playouts(pos, per_pos=256, seed=7, epoch=2, policy=policy)
    engine.set_tuning(block, bpc, refill, defer)
    try:
        other = engine.run_traced(pos, per_pos=256, seed=7, epoch=2, policy=policy)
        other_fast = engine.run_playouts(pos, per_pos=256, seed=7, epoch=2, policy=policy)
        assert other.stats["block_size"] == block and other.stats["refill"] == refill
    finally:
        engine.set_tuning()
    assert np.array_equal(base.winners, other.winners)
    assert np.array_equal(base.hashes, other.hashes)
    assert np.array_equal(fast.winners, base.winners)
    assert np.array_equal(other_fast.winners, base.winners)


def test_uneven_counts_empty_and_single(engine, oracle, golden):
    pos = golden["positions"]["open1024"][:40]
    counts = np.array([0, 1, 0, 0, 5, 63, 64, 65, 0, 1000] + [3] * 30, dtype=np.uint32)
    announce("uneven counts with zeros")
    res = engine.run_traced(pos, counts=counts, seed=5, epoch=0, policy=POLICY_RANDOM)
    _, tr = oracle.run_playouts(pos, counts=counts, seed=5, epoch=0, policy=POLICY_RANDOM)
    assert trace_mismatches(res, tr) == 0
    announce("empty batch")
    empty = engine.run_playouts(pos, counts=np.zeros(40, dtype=np.uint32))
    assert len(empty.winners) == 0 and empty.stats["playouts"] == 0
    none = engine.run_playouts(pos[:0], per_pos=10)
    assert len(none.winners) == 0
    announce("single playout")
    one = engine.run_traced(pos[3:4], per_pos=1, seed=11, epoch=4, policy=POLICY_HEURISTIC)
    _, tr1 = oracle.run_playouts(pos[3:4], per_pos=1, seed=11, epoch=4, policy=POLICY_HEURISTIC)
    assert trace_mismatches(one, tr1) == 0


@pytest.mark.parametrize("policy", [POLICY_RANDOM, POLICY_HEURISTIC, POLICY_HEURISTIC_NOISY])
def test_longest_first_order_uneven_counts(engine, oracle, golden, policy):
    """Batches large enough for the longest-first hand-out (at least 256 playouts per CU):
    uneven counts, zeros and repeated positions; every result slot and trace bit-exact."""
    pos = np.concatenate([golden["positions"]["open1024"], golden["positions"]["mid1024"]])
    rng = np.random.default_rng(17 + policy)
    counts = rng.integers(0, 700, len(pos)).astype(np.uint32)
    counts[::7] = 0
    announce(f"longest-first hand-out, uneven counts, {int(counts.sum())} playouts, policy {policy}")
    assert counts.sum() >= 2 * 256 * 1024
    res = engine.run_traced(pos, counts=counts, seed=21, epoch=3, policy=policy)
    _, tr = oracle.run_playouts(pos, counts=counts, seed=21, epoch=3, policy=policy)
    assert trace_mismatches(res, tr) == 0
    assert res.stats["plies"] == int(tr["plies"].sum(dtype=np.uint64))


@pytest.mark.parametrize("policy", [POLICY_RANDOM, POLICY_HEURISTIC_NOISY])
def test_ordered_batch_with_keys_and_tallies(engine, oracle, golden, policy):
    """A batch large enough for the longest-first hand-out, with stream keys, uneven counts and
    zeros: keyed winners against the oracle, and per-position tallies of the unkeyed batch
    against its winners (the tally kernel reads the offsets in position order while the
    playouts were handed out in rank order)."""
    pos = np.concatenate([golden["positions"]["open1024"], golden["positions"]["mid1024"]])
    rng = np.random.default_rng(29 + policy)
    counts = rng.integers(0, 96, len(pos)).astype(np.uint32)
    counts[::5] = 0
    keys = rng.integers(0, 2**32, len(pos), dtype=np.uint64).astype(np.uint32)
    total = int(counts.sum())
    assert total >= 256 * engine.device_info()["compute_units"]
    announce(f"ordered batch: keys and tallies, {total} playouts, policy {policy}")
    got = engine.run_playouts(pos, counts=counts, seed=13, epoch=6, policy=policy, keys=keys)
    w, _ = oracle.run_playouts(pos, counts=counts, seed=13, epoch=6, policy=policy, threads=16,
                               traces=False, keys=keys)
    assert np.array_equal(got.winners, w)
    plain = engine.run_playouts(pos, counts=counts, seed=13, epoch=6, policy=policy)
    assert not np.array_equal(plain.winners, got.winners)
    tallies, st = engine.run_tally(pos, counts=counts, seed=13, epoch=6, policy=policy)
    assert st["playouts"] == total
    offs = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])
    expect = np.zeros((len(pos), 3), dtype=np.uint32)
    for i in range(len(pos)):
        seg = plain.winners[offs[i]:offs[i + 1]]
        expect[i] = [(seg == -1).sum(), (seg == 0).sum(), (seg == 1).sum()]
    assert np.array_equal(tallies, expect)


def test_packed_batches_alternate_layouts(engine, oracle, golden):
    """Calls on one context whose packed batches differ in layout, one after another: ordered
    and in index order, with and without stream keys and uneven counts, larger and smaller than
    the last, with tallies and traces between them.  A section left over from an earlier call's
    layout would show up in some call's winners, traces or tallies against the oracle."""
    opens, mids = golden["positions"]["open1024"], golden["positions"]["mid1024"]
    rng = np.random.default_rng(41)
    keys = rng.integers(0, 2**32, 1024, dtype=np.uint64).astype(np.uint32)
    uneven = rng.integers(0, 200, 1024).astype(np.uint32)
    calls = [
        ("ordered, keyed, 1024 x 96", mids, dict(per_pos=96, keys=keys)),
        ("index order, 40 positions, uneven", opens[:40],
         dict(counts=uneven[:40] % 17)),
        ("ordered, uneven counts", opens, dict(counts=uneven)),
        ("index order, keyed, 8 x 3", mids[:8], dict(per_pos=3, keys=keys[:8])),
        ("ordered, 512 x 160", mids[:512], dict(per_pos=160)),
    ]
    for step, (name, pos, kw) in enumerate(calls):
        policy = (POLICY_RANDOM, POLICY_HEURISTIC, POLICY_HEURISTIC_NOISY)[step % 3]
        announce(f"packed batch {step}: {name}, policy {policy}")
        got = engine.run_playouts(pos, seed=19, epoch=step, policy=policy, **kw)
        w, _ = oracle.run_playouts(pos, seed=19, epoch=step, policy=policy, threads=16,
                                   traces=False, **kw)
        assert np.array_equal(got.winners, w), name
        if "keys" not in kw:
            tallies, _ = engine.run_tally(pos, seed=19, epoch=step, policy=policy, **kw)
            counts = kw["counts"] if "counts" in kw else np.full(len(pos), kw["per_pos"])
            offs = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])
            for i in range(0, len(pos), 37):
                seg = w[offs[i]:offs[i + 1]]
                assert list(tallies[i]) == [(seg == -1).sum(), (seg == 0).sum(),
                                            (seg == 1).sum()], (name, i)
        if step == 1:
            res = engine.run_traced(pos, seed=19, epoch=step, policy=policy, **kw)
            _, tr = oracle.run_playouts(pos, seed=19, epoch=step, policy=policy, **kw)
            assert trace_mismatches(res, tr) == 0


def test_batch_beyond_pinned_staging(engine, oracle, golden):
    """A batch whose winners exceed the 64 MiB of pinned staging (copied straight to the
    caller's memory) returns what the same positions give in a small, staged call."""
    pos = golden["positions"]["open1024"]
    per = 65600  # 1024 x 65,600 = 67.2M playouts > 64 MiB of winners
    announce(f"mi_run_playouts beyond pinned staging: {len(pos) * per} playouts")
    big = engine.run_playouts(pos, per_pos=per, seed=8, epoch=1)
    assert big.winners.size == len(pos) * per
    assert set(np.unique(big.winners)) <= {-1, 0, 1}
    # a position's stream is keyed by its index in the batch, so counts that are zero elsewhere
    # pick any slice of it: the first three, a middle block and the last three positions
    for lo, hi in ((0, 3), (500, 508), (1021, 1024)):
        counts = np.zeros(len(pos), dtype=np.uint32)
        counts[lo:hi] = per
        part = engine.run_playouts(pos, counts=counts, seed=8, epoch=1)
        assert np.array_equal(big.winners[lo * per:hi * per], part.winners), (lo, hi)
    # and one position of the batch against the oracle
    counts = np.zeros(len(pos), dtype=np.uint32)
    counts[777] = per
    w, _ = oracle.run_playouts(pos, counts=counts, seed=8, epoch=1, traces=False)
    assert np.array_equal(big.winners[777 * per:778 * per], w)


def test_direct_upload_beyond_pinned_staging(engine, oracle, golden):
    """A batch whose positions and offsets exceed the 64 MiB of pinned staging (about 3.4M
    positions, uploaded straight from the caller's memory): slices against small calls with zero
    counts elsewhere, and against the oracle."""
    base = golden["positions"]["open1024"]
    pos = np.tile(base, 4096)  # 4,194,304 positions x 20 bytes of position and offset
    counts = np.ones(len(pos), dtype=np.uint32)
    counts[1::2] = 2
    announce(f"mi_run_playouts with {len(pos)} positions (direct upload), "
             f"{int(counts.sum())} playouts")
    big = engine.run_playouts(pos, counts=counts, seed=3, epoch=5)
    offs = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])
    assert big.winners.size == offs[-1]
    for lo, hi in ((0, 4), (2_000_000, 2_000_006), (len(pos) - 4, len(pos))):
        part_counts = np.zeros(len(pos), dtype=np.uint32)
        part_counts[lo:hi] = counts[lo:hi]
        part = engine.run_playouts(pos, counts=part_counts, seed=3, epoch=5)
        assert np.array_equal(big.winners[offs[lo]:offs[hi]], part.winners), (lo, hi)
        w, _ = oracle.run_playouts(pos, counts=part_counts, seed=3, epoch=5, traces=False)
        assert np.array_equal(part.winners, w), (lo, hi)


def test_terminal_and_odd_positions(engine, oracle, golden):
    ml = golden["movelists"]["positions"]
    # random placements and king-loop positions sit at the end of the golden set; they include
    # finished games (counter >= 50, no moves) and men on rows a game never leaves them on
    odd = ml[-3300:]
    for policy in (POLICY_RANDOM, POLICY_HEURISTIC, POLICY_HEURISTIC_NOISY):
        announce(f"odd positions x 8, policy {policy}")
        res = engine.run_traced(odd, per_pos=8, seed=31, epoch=0, policy=policy)
        _, tr = oracle.run_playouts(odd, per_pos=8, seed=31, epoch=0, policy=policy)
        assert trace_mismatches(res, tr) == 0
    finished = odd[odd["counter"] >= 50]
    assert len(finished) > 0
    res = engine.run_traced(finished, per_pos=2, seed=1)
    assert np.all(res.winners == -1) and np.all(res.plies == 0)


def test_tally_staged_and_repeat(engine, golden):
    pos = golden["positions"]["open1024"][:300]
    counts = (np.arange(300, dtype=np.uint32) * 37) % 200
    announce("tally / staged / repeat")
    res = engine.run_playouts(pos, counts=counts, seed=77, epoch=3, policy=POLICY_HEURISTIC)
    tallies, _ = engine.run_tally(pos, counts=counts, seed=77, epoch=3, policy=POLICY_HEURISTIC)
    offs = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])
    for i in range(len(pos)):
        w = res.winners[offs[i]:offs[i + 1]]
        assert list(tallies[i]) == [int((w == -1).sum()), int((w == 0).sum()), int((w == 1).sum())]
    again = engine.run_playouts(pos, counts=counts, seed=77, epoch=3, policy=POLICY_HEURISTIC)
    assert np.array_equal(res.winners, again.winners)
    engine.stage(pos, counts=counts)
    engine.launch_staged(seed=77, epoch=3, policy=POLICY_HEURISTIC)
    assert np.array_equal(engine.fetch_winners(), res.winners)
    other = engine.run_playouts(pos, counts=counts, seed=77, epoch=4, policy=POLICY_HEURISTIC)
    assert not np.array_equal(other.winners, res.winners)


def test_device_start_positions_match_reference_golden(engine, golden):
    announce("mi_gen_start_positions for configs 0/1 and 2")
    open1024 = engine.start_positions(2016, 1024, 0, 30, False)
    mid1024 = engine.start_positions(2016, 1024, 20, 60, True)
    assert open1024.tobytes() == golden["positions"]["open1024"].tobytes()
    assert mid1024.tobytes() == golden["positions"]["mid1024"].tobytes()


def test_start_positions_reject_bad_ply_ranges(engine, oracle):
    """kmax < kmin, and a range of 2^32 ply numbers (kmax - kmin + 1 wraps to zero), are refused
    before anything runs.  Single ply counts and ranges of 2^31 ply numbers or more (the draw
    r mod (kmax - kmin + 1) with a divisor past 2^31) give the oracle's positions."""
    from gpu_ai_amd.engine import MI_ERR_INVALID_ARGUMENT, MiError
    announce("mi_gen_start_positions with bad and extreme ply ranges")
    for kmin, kmax in ((5, 4), (0, 2**32 - 1)):
        with pytest.raises(MiError) as e:
            engine.start_positions(3, 8, kmin, kmax)
        assert e.value.code == MI_ERR_INVALID_ARGUMENT
    for kmin, kmax in ((0, 0), (7, 7), (3, 2**31 + 8), (1, 2**32 - 1)):
        got = engine.start_positions(3, 256, kmin, kmax)  # games end long before k plies
        assert got.tobytes() == oracle.start_positions(3, 256, kmin, kmax).tobytes(), (kmin, kmax)


@pytest.mark.parametrize("refill", [REFILL_QUEUE, REFILL_STATIC])
@pytest.mark.parametrize("policy", [POLICY_RANDOM, POLICY_HEURISTIC, POLICY_HEURISTIC_NOISY])
def test_graph_replays_match_eager(engine, golden, policy, refill):
    """Replays one graph on batches of different sizes (each replay sizes the playout grid to its
    batch), with eager launches and synchronisations in between, and checks every replay against
    mi_run_playouts under both refill policies."""
    pos = golden["positions"]["mid1024" if policy else "open1024"]
    rng = np.random.default_rng(9 + policy + 2 * refill)
    engine.set_tuning(64, 12, refill, 16)
    try:
        with engine.graph(max_positions=1024, max_playouts=300000, policy=policy) as g:
            for rep in range(6):
                npos = int(rng.integers(1, 1024))
                sel = pos[rng.choice(len(pos), npos, replace=False)]
                # small and large batches alternate, so the grid shrinks and grows
                counts = rng.integers(0, 8 if rep % 2 else 400, npos).astype(np.uint32)
                announce(f"graph replay {rep}: {npos} positions, {int(counts.sum())} playouts, "
                         f"policy {policy}, refill {refill}")
                got = g.run(sel, counts, seed=31 + rep, epoch=rep)
                eager = engine.run_playouts(sel, counts=counts, seed=31 + rep, epoch=rep,
                                            policy=policy)
                assert np.array_equal(got.winners, eager.winners), rep
                assert got.stats["plies"] == eager.stats["plies"]
                if counts.sum():
                    assert got.stats["grid_blocks"] == eager.stats["grid_blocks"], rep
            with pytest.raises(Exception):
                g.run(pos[:2], np.array([200000, 200000], dtype=np.uint32))
    finally:
        engine.set_tuning(64, 12, REFILL_QUEUE, 16)


def test_valu_rate_probe(engine):
    """The measured int32 VALU issue rate (the roofline's peak) lies between a quarter of and the
    device's CUs x 64 lanes x reported clock; bad arguments are rejected."""
    info = engine.device_info()
    announce(f"mi_valu_rate on {info['compute_units']} CUs at {info['clock_khz']} kHz")
    rate, ms = engine.valu_rate(iters=4096, repeats=3)
    ceiling = info["compute_units"] * 64 * info["clock_khz"] * 1e3
    assert ms > 0.0
    assert 0.25 * ceiling < rate <= 1.02 * ceiling, (rate, ceiling)
    with pytest.raises(Exception):
        engine.valu_rate(iters=0)
    with pytest.raises(Exception):
        engine.valu_rate(iters=(1 << 20) + 1)


def test_autotune_picks_the_fastest_candidate(engine, golden):
    """mi_autotune measures every launch shape on the batch, keeps the fastest as the context's
    tuning, and the shape changes nothing but the schedule."""
    pos = golden["positions"]["open1024"][:128]
    before = engine.run_playouts(pos, per_pos=256, seed=21, epoch=2).winners
    announce("mi_autotune on 128 positions x 256 playouts")
    best, rows = engine.autotune(pos, per_pos=256, reps=1)
    assert len(rows) >= 100
    shapes = {(r["block_size"], r["blocks_per_cu"], r["refill"], r["defer_min"]) for r in rows}
    key = (best["block_size"], best["blocks_per_cu"], best["refill"], best["defer_min"])
    assert key in shapes
    def shape(r):
        return (r["block_size"], r["blocks_per_cu"], r["refill"], r["defer_min"])
    assert len(shapes) == len(rows)  # each shape measured once
    best_ms = next(r["kernel_ms"] for r in rows if shape(r) == key)
    assert best_ms == min(r["kernel_ms"] for r in rows)
    for r in rows:  # every candidate fills whole SIMDs with whole blocks
        assert (r["block_size"] * r["blocks_per_cu"]) % 256 == 0
    assert engine.tuning == best
    after = engine.run_playouts(pos, per_pos=256, seed=21, epoch=2).winners
    assert np.array_equal(before, after)
    engine.set_tuning(64, 12, REFILL_QUEUE, 16)  # back to the defaults for the tests that follow
