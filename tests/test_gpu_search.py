"""Searches on the device: the drop-in driver under GameTree, batched multi-search, run_ai."""
from __future__ import annotations

import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def test_multisearch_graph_equals_plain_batched():
    from gpu_ai_amd import host_api
    announce("multi-search 24 searches, graph vs batched")
    kw = dict(searches=24, min_budget=200, max_budget=3000, initial=64, scale=0.5)
    g = host_api.multisearch(mode="graph", **kw)
    b = host_api.multisearch(mode="batched", **kw)
    assert g["playouts"] == b["playouts"] > 0
    assert g["iterations"] == b["iterations"]
    assert g["root_digest"] == b["root_digest"]
    announce("multi-search per-search launches")
    p = host_api.multisearch(mode="per_search", **kw)
    assert p["playouts"] == g["playouts"]
    assert p["root_digest"] == g["root_digest"]  # keyed streams: the same trees


def test_multisearch_pipelined_is_deterministic():
    from gpu_ai_amd import host_api
    announce("multi-search pipelined, twice, and graph mode")
    kw = dict(searches=24, min_budget=200, max_budget=3000, initial=64, scale=0.5)
    a = host_api.multisearch(mode="pipelined", **kw)
    b = host_api.multisearch(mode="pipelined", **kw)
    g = host_api.multisearch(mode="graph", **kw)
    assert a["root_digest"] == b["root_digest"] == g["root_digest"]
    assert a["playouts"] == b["playouts"] == g["playouts"] > 0  # every budget is spent


def test_graph_launch_finish_on_two_contexts(golden):
    """Two contexts, two graphs in flight at once; each replay equals the plain call."""
    from gpu_ai_amd.engine import Engine, MiError
    pos = golden["positions"]["mid1024"][:300]
    counts = np.random.default_rng(3).integers(0, 40, len(pos)).astype(np.uint32)
    announce("two graphs in flight")
    with Engine(0) as e1, Engine(0) as e2:
        ref1 = e1.run_playouts(pos, counts=counts, seed=5, epoch=1).winners
        ref2 = e2.run_playouts(pos, counts=counts, seed=6, epoch=2).winners
        with e1.graph(len(pos), int(counts.sum())) as g1, e2.graph(len(pos), int(counts.sum())) as g2:
            with pytest.raises(MiError) as e:
                g1.finish()
            assert e.value.code == 6  # nothing launched
            g1.launch(pos, counts, seed=5, epoch=1)
            with pytest.raises(MiError):
                g1.launch(pos, counts, seed=5, epoch=1)  # one replay in flight per graph
            g2.launch(pos, counts, seed=6, epoch=2)
            r1 = g1.finish()
            r2 = g2.finish()
            assert np.array_equal(r1.winners, ref1) and np.array_equal(r2.winners, ref2)
            assert np.array_equal(g1.run(pos, counts, seed=5, epoch=1).winners, ref1)


def test_run_ai_modes():
    from gpu_ai_amd import host_api
    exe = host_api.run_ai_path()
    announce("run_ai gen_moves_test")
    out = subprocess.run([exe, "-m", "gen_moves_test", "-n", "3000"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0 and "Passed" in out.stdout, out.stdout
    announce("run_ai playout_test")
    out = subprocess.run([exe, "-m", "playout_test", "-n", "20000", "-1", "device_mi355x", "-2",
                          "device_coarse"], capture_output=True, text=True, timeout=120)
    # both names select the same engine and the same streams except for the epoch: results are
    # statistically alike, and the tallies are printed as the reference prints them
    assert out.returncode == 0 and "Similar results" in out.stdout, out.stdout
    announce("run_ai game_test random vs mcts_device_mi355x")
    out = subprocess.run([exe, "-m", "game_test", "-n", "2", "-1", "random", "-2",
                          "mcts_device_mi355x", "--timeout", "0.02"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0 and "=== Results ===" in out.stdout, out.stdout
    announce("run_ai unknown driver")
    out = subprocess.run([exe, "-m", "playout_test", "-n", "10", "-1", "nope"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 1 and "Unknown playout type" in out.stdout


def test_run_ai_net_leaf_player(tmp_path):
    """mcts_device_mi355x_net: the run_ai player whose leaves the value network scores."""
    import os
    from gpu_ai_amd import host_api
    from gpu_ai_amd.engine import Engine, ValueNet
    exe = host_api.run_ai_path()
    ckpt = str(tmp_path / "value.ckpt")
    announce("value network: a few training steps, checkpoint")
    with Engine(0) as eng:
        pos = eng.start_positions(3, 2048, 0, 60, reject_terminal=True)
        tallies, _ = eng.run_tally(pos, per_pos=32, seed=9)
        p1 = pos["turn"] == 0
        t = (np.where(p1, tallies[:, 1], tallies[:, 2]).astype(np.float32) -
             np.where(p1, tallies[:, 2], tallies[:, 1])) / 32.0
        with ValueNet(eng, 1024, seed=5) as net:
            for k in range(4):
                net.train_step(pos[k * 512:(k + 1) * 512], t[k * 512:(k + 1) * 512])
            net.save(ckpt)
    env = dict(os.environ, GPUAI_VALUE_NET=ckpt)
    announce("run_ai game_test mcts_device_mi355x_net vs random")
    out = subprocess.run([exe, "-m", "game_test", "-n", "2", "-1", "mcts_device_mi355x_net", "-2",
                          "random", "--timeout", "0.02"], capture_output=True, text=True,
                         timeout=300, env=env)
    assert out.returncode == 0 and "=== Results ===" in out.stdout, out.stdout + out.stderr
    announce("run_ai with a missing checkpoint")
    env["GPUAI_VALUE_NET"] = str(tmp_path / "missing.ckpt")
    out = subprocess.run([exe, "-m", "game_test", "-n", "1", "-1", "mcts_device_mi355x_net", "-2",
                          "random", "--timeout", "0.02"], capture_output=True, text=True,
                         timeout=120, env=env)
    assert out.returncode != 0 and "cannot load the value network" in out.stdout + out.stderr


def test_net_match_small():
    """The net-leaf vs playout-leaf match harness on a few short games: consistent tallies and
    per-part times, several exploration settings of both players, and rejected configurations."""
    from gpu_ai_amd import host_api
    from gpu_ai_amd.engine import Engine, ValueNet
    with Engine(0) as eng, ValueNet(eng, 4096, seed=7) as net:
        params = net.get_params()
    for k, kp in ((2.0, 2.0), (0.5, 8.0), (64.0, 0.5)):
        announce(f"net match: 4 games, 10 ms per move, 40 plies, exploration {k} / {kp}")
        r = host_api.net_match(params, games=4, move_ms=10.0, max_plies=40, net_exploration=k,
                               playout_exploration=kp)
        assert r["games"] == 4
        assert r["net_wins"] + r["playout_wins"] + r["draws"] == 4
        out = np.asarray(r["outcome"])
        assert set(out.tolist()) <= {-1, 0, 1}
        assert int(np.sum(out == 1)) == r["net_wins"] and int(np.sum(out == -1)) == r["playout_wins"]
        assert r["net_leaves"] > 0 and r["playout_leaves"] > 0 and r["plies"] > 0
        assert all(v >= 0.0 for key in ("select_ms", "score_ms", "update_ms") for v in r[key])
        assert r["score_ms"][0] > 0.0 and r["score_ms"][1] > 0.0
    with pytest.raises(RuntimeError):
        host_api.net_match(params, games=2, move_ms=5.0, max_plies=10, net_exploration=0.0)
    with pytest.raises(RuntimeError):
        host_api.net_match(params, games=2, move_ms=5.0, max_plies=10, playout_exploration=-1.0)


# ---------------------------------------------------------------------------------------------
# stream keys and the searches' batches against the oracle
# ---------------------------------------------------------------------------------------------

def _group_runs(states):
    """Runs of equal consecutive states -> (positions, counts), as runPlayouts groups them."""
    keyf = lambda p: (int(p["p1"]), int(p["p2"]), int(p["kings"]), int(p["turn"]),
                      int(p["counter"]))
    pos, counts = [], []
    for i, p in enumerate(states):
        if i > 0 and keyf(p) == keyf(states[i - 1]):
            counts[-1] += 1
        else:
            pos.append(p)
            counts.append(1)
    return np.array(pos, dtype=states.dtype), np.array(counts, dtype=np.uint32)


def test_keyed_playouts_match_oracle_and_ignore_batch_layout(engine, oracle, golden):
    """mi_run_playouts_keyed and keyed graph replays against the oracle; a keyed position plays
    the same games whatever shares its batch."""
    pos = golden["positions"]["mid1024"][:400]
    rng = np.random.default_rng(12)
    counts = rng.integers(0, 60, len(pos)).astype(np.uint32)
    keys = rng.integers(0, 2**32, len(pos), dtype=np.uint64).astype(np.uint32)
    for policy in (0, 1):
        announce(f"keyed playouts, policy {policy}")
        got = engine.run_playouts(pos, counts=counts, seed=71, epoch=9, policy=policy, keys=keys)
        w, _ = oracle.run_playouts(pos, counts=counts, seed=71, epoch=9, policy=policy,
                                   traces=False, keys=keys)
        assert np.array_equal(got.winners, w)
        # positions 100..199 alone, with their keys: the same winners as inside the batch
        offs = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])
        part = engine.run_playouts(pos[100:200], counts=counts[100:200], seed=71, epoch=9,
                                   policy=policy, keys=keys[100:200])
        assert np.array_equal(part.winners, got.winners[offs[100]:offs[200]])
        with engine.graph(len(pos), int(counts.sum()), policy=policy) as g:
            rep = g.run(pos, counts, seed=71, epoch=9, keys=keys)
            assert np.array_equal(rep.winners, w)
            plain = g.run(pos, counts, seed=71, epoch=9)  # unkeyed replay: key = index
            eager = engine.run_playouts(pos, counts=counts, seed=71, epoch=9, policy=policy)
            assert np.array_equal(plain.winners, eager.winners)


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_driver_run_playouts_matches_oracle(golden, oracle, policy):
    """DeviceMI355XPlayoutDriver::runPlayouts on a leaf vector with repeated consecutive states
    (including finished games, as terminal tree nodes hand out copies): each call's winners equal
    the oracle's on the grouped batch with the call's epoch."""
    from gpu_ai_amd import host_api
    mid = golden["positions"]["mid1024"][:60]
    odd = golden["movelists"]["positions"]
    finished = odd[odd["counter"] >= 50][:5]
    rng = np.random.default_rng(30 + policy)
    pool = np.concatenate([mid, finished])
    picks = rng.integers(0, len(pool), 150)
    reps = rng.integers(1, 6, 150)
    states = np.repeat(pool[picks], reps)  # runs of 1-5 equal states, equal runs may recur
    announce(f"driver runPlayouts on {len(states)} states, policy {policy}")
    got = host_api.driver_run_playouts(states, policy=policy, seed=4321, calls=2)
    gpos, gcounts = _group_runs(states)
    for c in range(2):
        w, _ = oracle.run_playouts(gpos, counts=gcounts, seed=4321, epoch=c, policy=policy,
                                   traces=False)
        assert np.array_equal(got[c], w), c


def test_search_leaves_match_oracle(golden, oracle):
    """A GameTree search driven by runPlayouts (MCTSPlayer::worker's loop): every round's leaves,
    grouped as the driver groups them, get the oracle's winners.  The second root is two plies
    from the draw limit, so its tree reaches finished games and hands out repeated states."""
    from gpu_ai_amd import host_api
    odd = golden["movelists"]["positions"]
    near_draw = next(p for p in odd[odd["counter"] == 48] if not oracle.is_game_over(p))
    roots = [golden["positions"]["mid1024"][7], near_draw]
    repeats = 0
    for k, root in enumerate(roots):
        announce(f"search rounds from root {k}")
        rounds = host_api.search_record(root, iterations=6, initial=300, scale=0.0, policy=0,
                                        seed=777)
        assert len(rounds) == 6
        for r in rounds:
            gpos, gcounts = _group_runs(r["pos"])
            repeats += int((gcounts > 1).sum())
            w, _ = oracle.run_playouts(gpos, counts=gcounts, seed=777, epoch=r["iteration"],
                                       policy=0, traces=False)
            assert np.array_equal(r["winners"], w), (k, r["iteration"])
    assert repeats > 0  # the grouping of repeated leaves was exercised


def test_multisearch_256_matches_oracle_and_modes_agree(oracle):
    """The batched multi-search at its configured size: 256 searches from seeded distinct
    positions, budgets uniform in [1000, 64000], graph replays of the persistent refill kernel.
    Every batch's winners equal the oracle's on the same positions, counts, keys and epoch; every
    budget is spent exactly; batched, pipelined and per-search launches grow the same trees."""
    from gpu_ai_amd import host_api
    stats = {"batches": 0, "playouts": 0, "bad": 0}
    per_search = np.zeros(256, dtype=np.int64)

    def check(b):
        w, _ = oracle.run_playouts(b["pos"], counts=b["counts"], seed=b["seed"],
                                   epoch=b["iteration"], policy=b["policy"], traces=False,
                                   keys=b["keys"], threads=16)
        stats["bad"] += int(np.sum(w != b["winners"]))
        stats["batches"] += 1
        stats["playouts"] += len(w)
        np.add.at(per_search, (b["keys"] >> 16).astype(np.int64), b["counts"].astype(np.int64))

    announce("multisearch-256 graph mode, every batch against the oracle")
    g = host_api.multisearch(mode="graph", on_batch=check)
    announce(f"{stats['batches']} batches, {stats['playouts']} playouts checked")
    assert stats["bad"] == 0
    assert stats["playouts"] == g["playouts"] == int(g["budgets"].sum())
    assert np.all(g["budgets"] >= 1000) and np.all(g["budgets"] <= 64000)
    assert np.array_equal(per_search, g["budgets"].astype(np.int64))
    assert np.array_equal(g["root_trials"], g["budgets"])
    for mode in ("batched", "pipelined", "per_search"):
        announce(f"multisearch-256 {mode} mode")
        r = host_api.multisearch(mode=mode, autotune=False)
        assert r["playouts"] == g["playouts"], mode
        assert np.array_equal(r["root_trials"], g["budgets"]), mode
        assert r["root_digest"] == g["root_digest"], mode
