"""ctypes bindings of the host layer's C entry points (include/gpuai/gpuai_host.h)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .build import HOST_LIB_PATH, RUN_AI_PATH
from .engine import MOVE_DTYPE, POS_DTYPE, library

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_lib = None

MODES = {"graph": 0, "batched": 1, "per_search": 2, "pipelined": 3}


class BatchRecord(ctypes.Structure):
    _fields_ = [("iteration", _U32), ("group", _U32), ("seed", _U32), ("policy", _U32),
                ("npos", _U32), ("pad", _U32), ("playouts", ctypes.c_uint64), ("pos", _P),
                ("counts", _P), ("keys", _P), ("winners", _P)]


BATCH_HOOK = ctypes.CFUNCTYPE(None, _P, ctypes.POINTER(BatchRecord))


def batch_from_record(rec: BatchRecord) -> dict:
    """Copies a batch record into numpy arrays (pos, counts and keys may be absent)."""
    n, total = rec.npos, rec.playouts

    def arr(ptr, dtype, count):
        if not ptr or count == 0:
            return None
        buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
        return np.frombuffer(bytes(buf), dtype=dtype).copy()

    return {"iteration": rec.iteration, "group": rec.group, "seed": rec.seed,
            "policy": rec.policy, "pos": arr(rec.pos, POS_DTYPE, n),
            "counts": arr(rec.counts, np.uint32, n), "keys": arr(rec.keys, np.uint32, n),
            "winners": arr(rec.winners, np.int8, total)}


class MultiSearchConfig(ctypes.Structure):
    _fields_ = [("searches", _U32), ("seed", _U32), ("playout_seed", _U32), ("min_budget", _U32),
                ("max_budget", _U32), ("initial", _U32), ("scale", ctypes.c_double),
                ("policy", _U32), ("mode", _U32), ("groups", _U32), ("autotune", _U32),
                ("hook", BATCH_HOOK), ("hook_user", _P), ("budgets", _P),
                ("root_trials", _P)]


class MultiSearchResult(ctypes.Structure):
    _fields_ = [("playouts", ctypes.c_uint64), ("iterations", _U32), ("pad", _U32),
                ("wall_seconds", ctypes.c_double), ("device_ms", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("playouts_per_second", ctypes.c_double),
                ("root_digest", ctypes.c_uint64), ("select_ms", ctypes.c_double),
                ("playout_ms", ctypes.c_double), ("update_ms", ctypes.c_double),
                ("block_size", _U32), ("blocks_per_cu", _U32), ("refill", _U32),
                ("defer_min", _U32)]


class NetMatchConfig(ctypes.Structure):
    _fields_ = [("games", _U32), ("seed", _U32), ("opening_min", _U32), ("opening_max", _U32),
                ("move_ms", ctypes.c_double), ("leaf_batch", _U32), ("max_plies", _U32),
                ("playout_seed", _U32), ("net_batch", _U32), ("progress_every", _U32),
                ("net_exploration", ctypes.c_double), ("playout_exploration", ctypes.c_double),
                ("mover_wins", _U32)]


class NetMatchResult(ctypes.Structure):
    _fields_ = [("games", _U32), ("net_wins", _U32), ("playout_wins", _U32), ("draws", _U32),
                ("score", ctypes.c_double), ("score_stderr", ctypes.c_double),
                ("net_leaves", ctypes.c_uint64), ("playout_leaves", ctypes.c_uint64),
                ("net_iterations", ctypes.c_uint64), ("playout_iterations", ctypes.c_uint64),
                ("plies", ctypes.c_uint64), ("seconds", ctypes.c_double),
                ("select_ms", ctypes.c_double * 2), ("score_ms", ctypes.c_double * 2),
                ("update_ms", ctypes.c_double * 2)]


def host_library() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError(f"{HOST_LIB_PATH} is missing: build it with `python -m gpu_ai_amd.build`")
        library()  # the device library it links against
        lib = ctypes.CDLL(HOST_LIB_PATH)
        lib.gpuai_host_gen_moves.argtypes = [_P, _P, _U32]
        lib.gpuai_host_gen_moves.restype = ctypes.c_int
        lib.gpuai_host_apply.argtypes = [_P, _P]
        lib.gpuai_host_apply.restype = ctypes.c_int
        lib.gpuai_mcts_trace.argtypes = [_P, _U32, _U32, ctypes.c_double, _P, _U32]
        lib.gpuai_mcts_trace.restype = ctypes.c_int
        lib.gpuai_multisearch_run.argtypes = [ctypes.POINTER(MultiSearchConfig),
                                              ctypes.POINTER(MultiSearchResult), ctypes.c_char_p,
                                              _U32]
        lib.gpuai_multisearch_run.restype = ctypes.c_int
        lib.gpuai_playout_policy.argtypes = [ctypes.c_char_p]
        lib.gpuai_playout_policy.restype = ctypes.c_int
        lib.gpuai_driver_run_playouts.argtypes = [_P, _U32, _U32, _U32, _U32, _P,
                                                  ctypes.c_char_p, _U32]
        lib.gpuai_driver_run_playouts.restype = ctypes.c_int
        lib.gpuai_search_record.argtypes = [_P, _U32, _U32, ctypes.c_double, _U32, _U32,
                                            BATCH_HOOK, _P, ctypes.c_char_p, _U32]
        lib.gpuai_search_record.restype = ctypes.c_int
        lib.gpuai_net_match_run.argtypes = [ctypes.POINTER(NetMatchConfig), _P,
                                            ctypes.POINTER(NetMatchResult), _P, ctypes.c_char_p,
                                            _U32]
        lib.gpuai_net_match_run.restype = ctypes.c_int
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def gen_moves(pos) -> np.ndarray:
    p = np.ascontiguousarray(pos, dtype=POS_DTYPE).reshape(1)
    out = np.zeros(128, dtype=MOVE_DTYPE)
    n = host_library().gpuai_host_gen_moves(_ptr(p), _ptr(out), 128)
    if n < 0:
        raise RuntimeError("move list longer than 128")
    return out[:n].copy()


def apply(pos, move) -> np.ndarray:
    p = np.array(pos, dtype=POS_DTYPE).reshape(1).copy()
    m = np.array(move, dtype=MOVE_DTYPE).reshape(1)
    if host_library().gpuai_host_apply(_ptr(p), _ptr(m)) != 0:
        raise ValueError("illegal move")
    return p[0]


def mcts_trace(root, iterations: int, initial: int, scale: float) -> np.ndarray:
    p = np.ascontiguousarray(root, dtype=POS_DTYPE).reshape(1)
    out = np.zeros(5 * iterations + 4, dtype=np.uint32)
    n = host_library().gpuai_mcts_trace(_ptr(p), iterations, initial, scale, _ptr(out), len(out))
    if n < 0:
        raise RuntimeError(f"gpuai_mcts_trace failed ({n})")
    return out[:n]


def run_ai_path() -> str:
    return RUN_AI_PATH


def multisearch(searches: int = 256, seed: int = 2016, playout_seed: int = 12345,
                min_budget: int = 1000, max_budget: int = 64000, initial: int = 256,
                scale: float = 0.0, policy: int = 0, mode: str = "graph",
                groups: int = 2, autotune: bool = True, on_batch=None) -> dict:
    """Batched multi-search: concurrent searches sharing the device; returns the run's
    statistics, each search's budget and its root's trials.  on_batch (optional) is called
    with every device batch as a dict of arrays (batch_from_record) before the trees update."""
    budgets = np.zeros(searches, dtype=np.uint32)
    root_trials = np.zeros(searches, dtype=np.uint32)
    failure = []

    def hook(_user, rec):
        try:
            on_batch(batch_from_record(rec.contents))
        except BaseException as e:  # reported after the search returns
            failure.append(e)

    cb = BATCH_HOOK(hook) if on_batch is not None else BATCH_HOOK()
    cfg = MultiSearchConfig(searches, seed, playout_seed, min_budget, max_budget, initial, scale,
                            policy, MODES[mode], groups, int(autotune), cb, None, _ptr(budgets),
                            _ptr(root_trials))
    res = MultiSearchResult()
    err = ctypes.create_string_buffer(512)
    if host_library().gpuai_multisearch_run(ctypes.byref(cfg), ctypes.byref(res), err, 512) != 0:
        raise RuntimeError(err.value.decode())
    if failure:
        raise failure[0]
    out = {k: getattr(res, k) for k, _ in MultiSearchResult._fields_ if k != "pad"}
    out["budgets"] = budgets
    out["root_trials"] = root_trials
    return out


def playout_policy(name: str) -> int:
    """The MI_POLICY_* that getPlayoutDriver(name) plays (-1: unknown name); needs no device."""
    return int(host_library().gpuai_playout_policy(name.encode()))


def driver_run_playouts(states, policy: int = 0, seed: int = 12345, calls: int = 1) -> np.ndarray:
    """DeviceMI355XPlayoutDriver::runPlayouts on `states`, `calls` times (call c draws epoch c);
    returns winners [calls, len(states)]."""
    st = np.ascontiguousarray(states, dtype=POS_DTYPE)
    out = np.zeros((calls, len(st)), dtype=np.int8)
    err = ctypes.create_string_buffer(512)
    if host_library().gpuai_driver_run_playouts(_ptr(st), len(st), policy, seed, calls, _ptr(out),
                                                err, 512) != 0:
        raise RuntimeError(err.value.decode())
    return out


def search_record(root, iterations: int, initial: int = 256, scale: float = 0.0,
                  policy: int = 0, seed: int = 12345) -> list[dict]:
    """One GameTree search driven by DeviceMI355XPlayoutDriver::runPlayouts; returns each
    round's leaves (in trial order) and the winners the driver gave them."""
    p = np.ascontiguousarray(root, dtype=POS_DTYPE).reshape(1)
    rounds, failure = [], []

    def hook(_user, rec):
        try:
            rounds.append(batch_from_record(rec.contents))
        except BaseException as e:
            failure.append(e)

    cb = BATCH_HOOK(hook)
    err = ctypes.create_string_buffer(512)
    if host_library().gpuai_search_record(_ptr(p), iterations, initial, scale, policy, seed, cb,
                                          None, err, 512) != 0:
        raise RuntimeError(err.value.decode())
    if failure:
        raise failure[0]
    return rounds


def net_match(params, games: int = 400, seed: int = 2016, opening_min: int = 0,
              opening_max: int = 6, move_ms: float = 100.0, leaf_batch: int = 4096,
              max_plies: int = 300, playout_seed: int = 12345, net_batch: int = 4096,
              progress_every: int = 0, net_exploration: float = 64.0,
              playout_exploration: float = 2.0, mover_wins: bool = False) -> dict:
    """Value-network leaves against playout leaves at equal wall time per move (netmatch.hpp);
    returns the tally, the network's score and the per-game outcomes (1 / -1 / 0).  The
    exploration factors are each player's k in sqrt(k ln N / n); mover_wins makes both trees
    count the chooser's own wins in UCB1 instead of the reference's child side to move."""
    w = np.ascontiguousarray(params, dtype=np.float32)
    cfg = NetMatchConfig(games, seed, opening_min, opening_max, move_ms, leaf_batch, max_plies,
                         playout_seed, net_batch, progress_every, net_exploration,
                         playout_exploration, int(mover_wins))
    res = NetMatchResult()
    outcome = np.zeros(games, dtype=np.int8)
    err = ctypes.create_string_buffer(512)
    if host_library().gpuai_net_match_run(ctypes.byref(cfg), _ptr(w), ctypes.byref(res),
                                          _ptr(outcome), err, 512) != 0:
        raise RuntimeError(err.value.decode())
    out = {k: getattr(res, k) for k, _ in NetMatchResult._fields_}
    for k in ("select_ms", "score_ms", "update_ms"):  # [network player, playout player]
        out[k] = [round(float(v), 1) for v in out[k]]
    out["outcome"] = outcome
    return out
