"""The C++ host layer (caller side of the drop-in): State/Move on bitboards, GameTree."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

from oracle import pyoracle


def test_host_gen_moves_golden(golden):
    from gpu_ai_amd import host_api
    ml = golden["movelists"]
    pos, counts, offs, moves = ml["positions"], ml["counts"], ml["offsets"], ml["moves"]
    for i in range(0, len(pos), 3):
        got = host_api.gen_moves(pos[i])
        ref = moves[int(offs[i]):int(offs[i]) + int(counts[i])]
        assert len(got) == counts[i] and got.tobytes() == ref.tobytes(), i


def test_host_apply_matches_oracle(oracle, golden):
    from gpu_ai_amd import host_api
    rng = np.random.default_rng(5)
    pos = golden["positions"]["mid1024"][:200]
    for p in pos:
        for _ in range(30):
            ms = oracle.gen_moves(p)
            if len(ms) == 0 or p["counter"] >= 50:
                break
            m = ms[rng.integers(len(ms))]
            a = host_api.apply(p, m)
            b = oracle.apply(p, m)
            assert a.tobytes() == b.tobytes()
            p = b


def test_host_apply_rejects_illegal(golden):
    from gpu_ai_amd import host_api
    p = golden["positions"]["open1024"][0]
    m = np.zeros(1, dtype=pyoracle.MOVE_DTYPE)[0]
    m["from"], m["to"] = 31, 0
    with pytest.raises(ValueError):
        host_api.apply(p, m)


@pytest.mark.skipif(not pyoracle.reference_available(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("which,iters,initial,scale", [("start", 40, 50, 0.02), ("mid1024_0", 30, 200, 0.25),
                                                      ("mid1024_5", 12, 64, 1.0), ("open1024_9", 60, 7, 0.1)])
def test_mcts_matches_reference_gametree(golden, which, iters, initial, scale):
    """GameTree select/update/UCB1/getOptMove against the reference's own src/mcts.cpp."""
    from gpu_ai_amd import host_api
    if which == "start":
        root = pyoracle.start_position()
    else:
        name, idx = which.split("_")
        root = golden["positions"][name][int(idx)]
    ref = pyoracle.Reference().mcts_trace(root, iters, initial, scale)
    mine = host_api.mcts_trace(root, iters, initial, scale)
    assert len(mine) == len(ref)
    assert np.array_equal(mine, ref)
    assert mine[5 * (iters - 1) + 2] > initial  # the tree grew


def test_run_ai_usage():
    from gpu_ai_amd import host_api
    out = subprocess.run([host_api.run_ai_path(), "-h"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "gen_moves_test" in out.stdout


def test_driver_names_select_the_reference_policies():
    """getPlayoutDriver's names (src/playout.cpp:189-223): the reference's heuristic drivers
    play the noisy heuristic (src/heuristicPlayout.cpp:33), hybrid_heuristic is a plain hybrid
    driver there (src/playout.cpp:211-213), names it does not know are rejected."""
    from gpu_ai_amd import host_api
    from gpu_ai_amd.engine import POLICY_HEURISTIC, POLICY_HEURISTIC_NOISY, POLICY_RANDOM
    expect = {"host": POLICY_RANDOM, "device_single": POLICY_RANDOM,
              "device_multiple": POLICY_RANDOM, "device_coarse": POLICY_RANDOM,
              "hybrid": POLICY_RANDOM, "hybrid_heuristic": POLICY_RANDOM,
              "optimal": POLICY_RANDOM, "device_mi355x": POLICY_RANDOM,
              "host_heuristic": POLICY_HEURISTIC_NOISY, "device_heuristic": POLICY_HEURISTIC_NOISY,
              "optimal_heuristic": POLICY_HEURISTIC_NOISY,
              "device_mi355x_heuristic": POLICY_HEURISTIC_NOISY,
              "device_mi355x_heuristic_argmax": POLICY_HEURISTIC,
              "hybrid_heruistic": -1, "": -1, "Host": -1}
    for name, policy in expect.items():
        assert host_api.playout_policy(name) == policy, name
