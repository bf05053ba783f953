"""Shared fixtures.  Tests marked `gpu` need an MI355X and the built device library."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.join(ROOT, "tests")
if TESTS not in sys.path:  # test helpers (devmem)
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")


def _ensure_oracle() -> None:
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    from oracle import pyoracle
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def golden():
    def load(name):
        with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
            return {k: z[k] for k in z.files}
    g = {
        "positions": load("positions.npz"),
        "movelists": load("movelists.npz"),
        "playouts": load("playouts.npz"),
        "scores": load("scores.npz")["scores"],
        "noisy": load("noisy.npz"),
    }
    with open(os.path.join(GOLDEN, "perft.json")) as f:
        g["perft"] = json.load(f)
    full = load("full_configs.npz")
    g["full"] = {name: {"winners": unpack2(full[name + "_winners2"]),
                        "plies": full[name + "_plies"]}
                 for name in ("random_1024x1024", "heuristic_1024x1024")}
    return g


def unpack2(packed: np.ndarray) -> np.ndarray:
    """Inverse of make_golden.pack2: four two-bit codes (winner + 1) per byte, low bits first."""
    p = packed.astype(np.uint8)
    codes = np.stack([(p >> s) & 3 for s in (0, 2, 4, 6)], axis=1).reshape(-1)
    return codes.astype(np.int8) - 1


@pytest.fixture(scope="session")
def engine():
    from gpu_ai_amd.engine import Engine
    eng = Engine(0)
    yield eng
    eng.close()
