"""Build recipe: compiles the gfx950 library (and the host programs) in-tree with hipcc.

    python -m gpu_ai_amd.build          # library + host programs
    python -m gpu_ai_amd.build --clean

Outputs next to this file: libgpuai_mi355x.so (C ABI of include/gpuai/mi_playout.h) and the
host programs under bin/.  Every HIP translation unit is compiled with packed-FP32 instructions
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(PKG, "_build")
BIN = os.path.join(PKG, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

LIB_NAME = "libgpuai_mi355x.so"
LIB_PATH = os.path.join(PKG, LIB_NAME)

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I" + os.path.join(ROOT, "include"),
          "-I" + CSRC, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function"]

# (source, object) pairs of the device library
LIB_SOURCES = ["playout_kernels.hip", "capi.cpp", "noalloc.hip", "mlp_kernels.hip", "mlp.cpp"]
# the C++ host layer (caller side: State, PlayoutDriver, GameTree, players, CLI)
HOST_SOURCES = ["state.cpp", "playout.cpp", "mcts.cpp", "player.cpp", "multisearch.cpp",
                "netleaf.cpp", "netmatch.cpp", "capi_host.cpp"]
HOST_LIB_PATH = os.path.join(PKG, "libgpuai_host.so")
RUN_AI_PATH = os.path.join(BIN, "run_ai")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; that line is noise
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def _compile(src: str, obj: str, extra: list[str] | None = None) -> str:
    cmd = [HIPCC] + COMMON + (extra or []) + ["-c", src, "-o", obj]
    _run(cmd)
    return obj


def build_library(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    pairs = [(os.path.join(CSRC, s), os.path.join(OBJ, s.replace(".", "_") + ".o")) for s in LIB_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda p: _compile(*p), pairs))
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_host(jobs: int = 8) -> None:
    """Builds libgpuai_host.so and bin/run_ai over the device library."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    extra = ["-fopenmp", "-I" + HOST]
    srcs = HOST_SOURCES + ["run_ai.cpp"]
    pairs = [(os.path.join(HOST, s), os.path.join(OBJ, "host_" + s.replace(".", "_") + ".o"))
             for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda p: _compile(p[0], p[1], extra), pairs))
    link = ["-fopenmp", "-L" + PKG, "-lgpuai_mi355x", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/.."]
    tmp = HOST_LIB_PATH + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs[:-1], "-o", tmp] + link)
    os.replace(tmp, HOST_LIB_PATH)
    tmp = RUN_AI_PATH + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", objs[-1], "-o", tmp, "-L" + PKG, "-lgpuai_host"] + link)
    os.replace(tmp, RUN_AI_PATH)


def build_oracle() -> None:
    """Builds oracle/liboracle.so and, where the reference sources exist, oracle/_ref."""
    oracle = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", oracle, "-B", "all"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-j8", "-C", oracle, "-B", "ref"])


def build_all(jobs: int = 8) -> None:
    build_library(jobs)
    build_host(jobs)
    build_oracle()


def clean() -> None:
    for p in (OBJ, BIN):
        shutil.rmtree(p, ignore_errors=True)
    for p in (LIB_PATH, HOST_LIB_PATH):
        if os.path.exists(p):
            os.remove(p)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    args = ap.parse_args()
    if args.clean:
        clean()
        return 0
    build_all(args.j)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
