"""The playout kernels' per-lane game logic, compiled for the host and checked against the oracle.

gpu_ai_amd/csrc/checkers.h is the code every playout kernel runs per lane (bitboard move
generation in the mover frame, slot selection, deferred multi-jump walks, heuristic weight
classes).  Its functions are __host__ __device__, so the same source builds for the CPU: the
checker (tests/lane_logic_check.cpp) plays kernel-shaped playouts on the host and compares them
with the oracle's traces, and compares move lists entry by entry.  The device's exact-division
path, the only part that differs between host and device builds, is covered by the GPU test
test_divmod_exact.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert proc.returncode == 0, proc.stdout


@pytest.fixture(scope="module")
def checker_obj(tmp_path_factory, oracle):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    obj = str(tmp_path_factory.mktemp("lane") / "lane_logic_check.o")
    _run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-host-only",
          "-I" + os.path.join(ROOT, "gpu_ai_amd", "csrc"), "-I" + os.path.join(ROOT, "include"),
          "-I" + os.path.join(ROOT, "oracle"), "-c",
          os.path.join(ROOT, "tests", "lane_logic_check.cpp"), "-o", obj])
    return obj


def _link(obj: str, oracle_so: str, out: str) -> str:
    _run([HIPCC, obj, oracle_so, "-Wl,-rpath," + os.path.dirname(oracle_so), "-o", out])
    return out


@pytest.fixture(scope="module")
def checker(checker_obj):
    return _link(checker_obj, os.path.join(ROOT, "oracle", "liboracle.so"),
                 checker_obj[:-2])


def test_lane_logic_matches_oracle(checker):
    proc = subprocess.run([checker, "512", "8"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=600)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout
    assert "mismatches 0" in proc.stdout
    # the sample reaches the deferred multi-jump path
    deferred = int(proc.stdout.split("deferred plies ")[1].split(")")[0])
    assert deferred > 1000
    # every move index of the multi-jump positions met, chosen the deferred ply's way
    choices = int(proc.stdout.split("deferred choices ")[1].split(" ")[0])
    assert choices > 10000


@pytest.mark.parametrize("k", range(1, 10))
def test_planted_oracle_mistake_is_caught(checker_obj, tmp_path, k):
    """The oracle rebuilt with one planted mistake (OR_MUTATION k in oracle/oracle.c: jump
    order, king relanding, draw ply, promotion in captures, heuristic ties, Philox rounds, the
    noisy heuristic's sigma, its draws indexed in reference order) must disagree with the lane
    logic."""
    so = str(tmp_path / f"liboracle_m{k}.so")
    _run(["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared", f"-DOR_MUTATION={k}",
          os.path.join(ROOT, "oracle", "oracle.c"), "-o", so])
    exe = _link(checker_obj, so, str(tmp_path / "check"))
    proc = subprocess.run([exe, "256", "8"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=600)
    assert proc.returncode != 0 and "mismatches 0" not in proc.stdout, proc.stdout
