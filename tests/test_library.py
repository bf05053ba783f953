"""The C-ABI library loads and exports every symbol its headers declare (no GPU needed)."""
from __future__ import annotations

import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(prefix: str) -> list[str]:
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "gpuai", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w \*]*?\b(" + prefix + r"\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared("mi_")
    for n in ("mi_create", "mi_destroy", "mi_run_playouts", "mi_gen_moves", "mi_strerror"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from gpu_ai_amd.build import LIB_PATH
    assert os.path.exists(LIB_PATH), "build the library first: python -m gpu_ai_amd.build"
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in declared("mi_") if not hasattr(lib, n)]
    assert not missing, missing


def test_host_library_exports_every_declared_symbol():
    from gpu_ai_amd.build import HOST_LIB_PATH
    assert os.path.exists(HOST_LIB_PATH), "build the host library first: python -m gpu_ai_amd.build"
    lib = ctypes.CDLL(HOST_LIB_PATH)
    names = declared("gpuai_")
    assert "gpuai_playout_policy" in names and "gpuai_driver_run_playouts" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_strerror_without_device():
    from gpu_ai_amd import engine
    lib = engine.library()
    assert lib.mi_strerror(0) == b"ok"
    assert lib.mi_strerror(engine.MI_ERR_CAPACITY).startswith(b"move list")
    assert lib.mi_strerror(999) == b"unknown error"


def test_null_context_is_rejected():
    from gpu_ai_amd import engine
    lib = engine.library()
    assert lib.mi_run_playouts(None, None, 0, None, 0, 0, 0, 0, None, None) == engine.MI_ERR_INVALID_ARGUMENT
    assert lib.mi_destroy(None) == engine.MI_OK


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_device_fails_cleanly():
    from gpu_ai_amd import engine
    with pytest.raises(engine.MiError) as e:
        engine.Engine(0)
    assert e.value.code == engine.MI_ERR_NO_DEVICE
