"""bench.py's contract pieces that need no GPU: the one-GPU refusal, the traffic figures it reads
from the rocprof summaries, and the training step's algorithmic FLOP count."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    proc = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert proc.returncode == 2
    assert "one GPU" in proc.stderr and proc.stdout == ""


@pytest.mark.parametrize("workload", ["random", "heuristic", "noisy", "small"])
def test_traffic_summaries_parse(workload):
    """Every playout workload has a traffic summary whose bytes per launch cover at least the
    algorithmic bytes: the positions read and one int8 winner written per playout."""
    cfg = bench.WORKLOADS[workload]
    traffic = bench.read_traffic(workload)
    assert traffic is not None
    algorithmic = cfg["n"] * 16 + cfg["n"] * cfg["per_pos"]
    assert algorithmic < traffic < 8 * algorithmic


def test_training_step_flops():
    """Forward 128*256 + 256*256 + 256 multiply-adds a position; backward the weight gradients of
    all three layers plus the input gradients of the two upper ones: 262,912 in all, 2.154 GFLOP
    for a batch of 4096."""
    fwd = 128 * 256 + 256 * 256 + 256
    bwd = (128 * 256 + 256 * 256 + 256) + (256 * 256 + 256)
    assert fwd + bwd == 262_912
    assert bench.mlp_step_flops(4096) == 2 * 4096 * 262_912 == 2_153_775_104
    assert bench.mlp_step_flops(1) * 4096 == bench.mlp_step_flops(4096)
