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


def test_multisearch_pipelined_is_deterministic():
    from gpu_ai_amd import host_api
    announce("multi-search pipelined, twice, and graph mode")
    kw = dict(searches=24, min_budget=200, max_budget=3000, initial=64, scale=0.5)
    a = host_api.multisearch(mode="pipelined", **kw)
    b = host_api.multisearch(mode="pipelined", **kw)
    g = host_api.multisearch(mode="graph", **kw)
    assert a["root_digest"] == b["root_digest"]
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
    per-part times, both exploration settings, and a rejected configuration."""
    from gpu_ai_amd import host_api
    from gpu_ai_amd.engine import Engine, ValueNet
    with Engine(0) as eng, ValueNet(eng, 4096, seed=7) as net:
        params = net.get_params()
    for k in (2.0, 0.5):
        announce(f"net match: 4 games, 10 ms per move, 40 plies, exploration {k}")
        r = host_api.net_match(params, games=4, move_ms=10.0, max_plies=40, net_exploration=k)
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
