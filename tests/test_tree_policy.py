"""Trial allocation of the batch UCB1 tree (gpu_ai_amd/host/mcts.cpp) with one root move rigged
to win, for the reference's exploitation term and for the mover's own (tests/tree_policy_check.cpp).

The reference's UCB1 counts the wins of the side to move in the child (src/mcts.cpp:189): the
opponent of the player choosing.  Its batch allocation (shares proportional to the UCB1 weights,
src/mcts.cpp:106-138) then gives the winning move the fewest trials; with the mover's term it
gets the most.  Both trees still play the winning move (getOptMove scores by the searching
player, src/mcts.cpp:39-55).
"""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = [os.path.join(ROOT, "tests", "tree_policy_check.cpp"),
           os.path.join(ROOT, "gpu_ai_amd", "host", "mcts.cpp"),
           os.path.join(ROOT, "gpu_ai_amd", "host", "state.cpp")]


def build(out: str, extra=()) -> str:
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    cmd = [HIPCC, "-O1", "-g", "-std=c++17", "-x", "hip", "--offload-host-only", "-fopenmp",
           "-I" + os.path.join(ROOT, "gpu_ai_amd", "host"),
           "-I" + os.path.join(ROOT, "gpu_ai_amd", "csrc"),
           "-I" + os.path.join(ROOT, "include"), *extra, *SOURCES, "-o", out]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert proc.returncode == 0, proc.stdout
    return out


def allocation(exe: str, *args) -> dict:
    proc = subprocess.run([exe, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=300)
    assert proc.returncode == 0, proc.stdout
    out: dict = {}
    for line in proc.stdout.splitlines():
        f = line.split()
        if len(f) == 5 and f[1] == "child":
            out.setdefault(f[0], {}).setdefault("trials", []).append(int(f[4]))
        elif len(f) == 3 and f[1] == "opt":
            out.setdefault(f[0], {})["opt"] = int(f[2])
    return out


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    return build(str(tmp_path_factory.mktemp("tree") / "tree_policy_check"))


@pytest.mark.parametrize("k,parallel", [(2.0, 0), (0.5, 0), (2.0, 1)])
def test_rigged_move_allocation(checker, k, parallel):
    a = allocation(checker, 200, 64, k, parallel)
    ref, mover = a["reference"]["trials"], a["mover"]["trials"]
    assert len(ref) == len(mover) == 7  # the starting position's moves
    assert sum(ref) == sum(mover) == 200 * 64
    assert ref[0] < min(ref[1:])      # the reference's term starves the winning move
    assert mover[0] > max(mover[1:])  # the mover's term favours it
    assert a["reference"]["opt"] == 0 and a["mover"]["opt"] == 0
