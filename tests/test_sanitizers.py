"""The host code under AddressSanitizer + UndefinedBehaviorSanitizer and ThreadSanitizer (CPU).

Built here from the sources:
  * tests/sanitize_check.cpp with the search code (gpu_ai_amd/host/player.cpp, mcts.cpp,
    state.cpp; netleaf.cpp and playout.cpp for the link): MCTSPlayer's worker thread searching
    while the game thread calls getMove and move under the tree mutex, and concurrent GameTree
    searches with OpenMP selection at the root, leaves scored by a host-only driver;
  * tests/lane_logic_check.cpp (the kernels' per-lane game logic, compiled for the host, against
    the oracle) and tests/tree_policy_check.cpp under ASan + UBSan;
  * the oracle's C restatement driven by tests/oracle_sanitize_check.c, built by gcc under
    ASan + UBSan and plainly, with the same digest of results required from both.
UBSan reports are fatal (-fno-sanitize-recover).  The OpenMP runtime is not instrumented: its
internal races are excluded (ignore_noninstrumented_modules) and the one allocation libomp keeps
to the end of the process is a suppressed leak.  Sanitizer flags go to the host compilation only.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HOST = os.path.join(ROOT, "gpu_ai_amd", "host")
LIB_DIR = os.path.join(ROOT, "gpu_ai_amd")
ASAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
        "-Xarch_host", "-fno-sanitize-recover=undefined"]
ASAN_LINK = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]
TSAN = ["-Xarch_host", "-fsanitize=thread"]


def _hipcc():
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    return HIPCC


def _build(out: str, sources: list[str], flags: list[str], link: list[str]) -> str:
    cmd = [_hipcc(), "-O1", "-g", "-std=c++17", "-x", "hip", "--offload-host-only", "-fopenmp",
           "-I" + HOST, "-I" + os.path.join(ROOT, "gpu_ai_amd", "csrc"),
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle"), *flags,
           *sources, "-x", "none", *link, "-o", out]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert proc.returncode == 0, proc.stdout
    return out


def _run(args: list[str], env_extra: dict, timeout: int = 600) -> str:
    env = dict(os.environ, **env_extra)
    proc = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=timeout, env=env)
    out = proc.stdout
    assert proc.returncode == 0, out[-4000:]
    for bad in ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:",
                "WARNING: ThreadSanitizer"):
        assert bad not in out, out[-4000:]
    return out


SEARCH_SOURCES = [os.path.join(ROOT, "tests", "sanitize_check.cpp")] + [
    os.path.join(HOST, f) for f in ("player.cpp", "mcts.cpp", "state.cpp", "netleaf.cpp",
                                    "playout.cpp")]
SEARCH_LINK = ["-L" + LIB_DIR, "-lgpuai_mi355x", "-Wl,-rpath," + LIB_DIR]


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    if not os.path.exists(os.path.join(LIB_DIR, "libgpuai_mi355x.so")):
        pytest.skip("the device library is not built")
    d = tmp_path_factory.mktemp("sanitize")
    (d / "lsan.supp").write_text("leak:___kmp_allocate\nleak:libomp.so\n")
    return d


def _asan_env(d) -> dict:
    return {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0",
            "LSAN_OPTIONS": f"suppressions={d / 'lsan.supp'}",
            "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}


@pytest.fixture(scope="module")
def search_asan(workdir):
    return _build(str(workdir / "search_asan"), SEARCH_SOURCES, ASAN, SEARCH_LINK + ASAN_LINK)


@pytest.fixture(scope="module")
def search_tsan(workdir):
    return _build(str(workdir / "search_tsan"), SEARCH_SOURCES, TSAN, SEARCH_LINK + TSAN)


@pytest.mark.parametrize("mode", ["tree", "player"])
def test_search_under_asan_ubsan(workdir, search_asan, mode):
    out = _run([search_asan, mode], _asan_env(workdir))
    assert ("player:" in out) if mode == "player" else ("tree 3: 5760 trials" in out)


@pytest.mark.parametrize("mode", ["tree", "player"])
def test_search_under_tsan(search_tsan, mode):
    out = _run([search_tsan, mode], {"TSAN_OPTIONS": "ignore_noninstrumented_modules=1:halt_on_error=1"})
    assert ("player:" in out) if mode == "player" else ("tree 3: 5760 trials" in out)


def test_tree_policy_under_asan_ubsan(workdir):
    exe = _build(str(workdir / "tree_policy_asan"),
                 [os.path.join(ROOT, "tests", "tree_policy_check.cpp"),
                  os.path.join(HOST, "mcts.cpp"), os.path.join(HOST, "state.cpp")],
                 ASAN, ASAN_LINK)
    out = _run([exe, "100", "64", "2", "1"], _asan_env(workdir))
    assert "mover opt 0" in out


def test_lane_logic_under_asan_ubsan(workdir, oracle):
    oracle_so = os.path.join(ROOT, "oracle", "liboracle.so")
    exe = _build(str(workdir / "lane_asan"), [os.path.join(ROOT, "tests", "lane_logic_check.cpp")],
                 ASAN, [oracle_so, "-Wl,-rpath," + os.path.dirname(oracle_so)] + ASAN_LINK)
    out = _run([exe, "96", "4"], _asan_env(workdir))
    assert "mismatches 0" in out


def test_oracle_under_asan_ubsan(tmp_path):
    """The oracle's C restatement (every entry point: start positions, move lists and moves,
    scores, playouts of each policy with uniform and uneven counts and stream keys, choice
    frequencies, noise draws) under AddressSanitizer + UndefinedBehaviorSanitizer, built with gcc,
    and the same digest of its results as a plain -O2 build."""
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("gcc not available")
    src = [os.path.join(ROOT, "oracle", "oracle.c"),
           os.path.join(ROOT, "tests", "oracle_sanitize_check.c")]
    common = ["-std=c11", "-fopenmp", "-I" + os.path.join(ROOT, "oracle")]
    digests = []
    for name, flags in (("plain", ["-O2"]),
                        ("sanitized", ["-O1", "-g", "-fno-omit-frame-pointer",
                                       "-fsanitize=address,undefined",
                                       "-fno-sanitize-recover=all"])):
        exe = str(tmp_path / name)
        proc = subprocess.run([gcc, *common, *flags, *src, "-lm", "-o", exe],
                              stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert proc.returncode == 0, proc.stdout
        out = _run([exe], {"ASAN_OPTIONS": "detect_leaks=1",
                           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}, timeout=300)
        digests.append(out.strip().splitlines()[-1])
    assert digests[0].startswith("digest ") and digests[0] == digests[1], digests
