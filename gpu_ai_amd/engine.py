"""ctypes bindings of the C ABI in include/gpuai/mi_playout.h and include/gpuai/mi_mlp.h.

The library must have been built (python -m gpu_ai_amd.build); importing this module on a box
without it raises at load time, so a missing extension never degrades into a silent fallback.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from dataclasses import dataclass

import numpy as np

from .build import LIB_PATH

POS_DTYPE = np.dtype([("p1", "<u4"), ("p2", "<u4"), ("kings", "<u4"), ("turn", "u1"),
                      ("counter", "u1"), ("pad", "u1", (2,))])
MOVE_DTYPE = np.dtype([("from", "u1"), ("to", "u1"), ("jumps", "u1"), ("promoted", "u1"),
                       ("inter", "u1", (8,)), ("removed", "u1", (8,))])
assert POS_DTYPE.itemsize == 16 and MOVE_DTYPE.itemsize == 20

MI_OK = 0
MI_ERR_INVALID_ARGUMENT = 1
MI_ERR_NO_DEVICE = 2
MI_ERR_DEVICE = 3
MI_ERR_OUT_OF_MEMORY = 4
MI_ERR_CAPACITY = 5
MI_ERR_NOT_STAGED = 6

POLICY_RANDOM = 0
POLICY_HEURISTIC = 1        # the heuristic without noise, ties broken by the ply's word
POLICY_HEURISTIC_NOISY = 2  # the reference's heuristic: an N(0, 0.11) draw on each weight
REFILL_STATIC = 0
REFILL_QUEUE = 1


class MiTuning(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("refill", ctypes.c_uint32), ("defer_min", ctypes.c_uint32)]


class MiTuneResult(ctypes.Structure):
    _fields_ = [("tuning", MiTuning), ("kernel_ms", ctypes.c_float),
                ("playouts_per_s", ctypes.c_float)]


class MiTuneEntry(ctypes.Structure):
    _fields_ = [("policy", ctypes.c_uint32), ("bucket", ctypes.c_uint32), ("tuning", MiTuning),
                ("kernel_ms", ctypes.c_float), ("playouts", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        t = self.tuning
        return {"policy": self.policy, "bucket": self.bucket, "block_size": t.block_size,
                "blocks_per_cu": t.blocks_per_cu, "refill": t.refill, "defer_min": t.defer_min,
                "kernel_ms": self.kernel_ms, "playouts": self.playouts}


class MiStats(ctypes.Structure):
    _fields_ = [("playouts", ctypes.c_uint64), ("plies", ctypes.c_uint64),
                ("kernel_ms", ctypes.c_float), ("total_ms", ctypes.c_float),
                ("max_moves", ctypes.c_uint32), ("grid_blocks", ctypes.c_uint32),
                ("block_size", ctypes.c_uint32), ("refill", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class MiError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_SIGS = {
    "mi_create": ([ctypes.POINTER(_P), ctypes.c_int, ctypes.POINTER(MiTuning)], ctypes.c_int),
    "mi_destroy": ([_P], ctypes.c_int),
    "mi_strerror": ([ctypes.c_int], ctypes.c_char_p),
    "mi_last_error": ([_P], ctypes.c_char_p),
    "mi_set_tuning": ([_P, ctypes.POINTER(MiTuning)], ctypes.c_int),
    "mi_get_tuning": ([_P, ctypes.POINTER(MiTuning)], ctypes.c_int),
    "mi_device_info": ([_P, _P, _P, _P], ctypes.c_int),
    "mi_run_playouts": ([_P, _P, _U32, _P, _U32, _U32, _U32, _U32, _P, ctypes.POINTER(MiStats)],
                        ctypes.c_int),
    "mi_run_playouts_keyed": ([_P, _P, _U32, _P, _U32, _P, _U32, _U32, _U32, _P,
                               ctypes.POINTER(MiStats)], ctypes.c_int),
    "mi_run_playouts_traced": ([_P, _P, _U32, _P, _U32, _U32, _U32, _U32, _P, _P, _P,
                                ctypes.POINTER(MiStats)], ctypes.c_int),
    "mi_run_playouts_tally": ([_P, _P, _U32, _P, _U32, _U32, _U32, _U32, _P,
                               ctypes.POINTER(MiStats)], ctypes.c_int),
    "mi_run_playouts_noalloc": ([_P, _P, _U32, _P, _U32, _U32, _U32, _U32, _P, _P, _P],
                                ctypes.c_int),
    "mi_gen_moves": ([_P, _P, _U32, _P, _U32, _P], ctypes.c_int),
    "mi_stage": ([_P, _P, _U32, _P, _U32], ctypes.c_int),
    "mi_launch_staged": ([_P, _U32, _U32, _U32, ctypes.POINTER(MiStats)], ctypes.c_int),
    "mi_fetch_winners": ([_P, _P, ctypes.c_uint64], ctypes.c_int),
    "mi_check_divmod": ([_P, _P, _P, _U32, _P, _P], ctypes.c_int),
    "mi_gen_start_positions": ([_P, _U32, _U32, _U32, _U32, _U32, _P], ctypes.c_int),
    "mi_autotune": ([_P, _P, _U32, _P, _U32, _U32, _U32, _P, _U32, _P, ctypes.POINTER(MiTuning)],
                    ctypes.c_int),
    "mi_autotune_sizes": ([_P, _P, _U32, _U32, _P, _U32, _U32, _P], ctypes.c_int),
    "mi_get_tuning_table": ([_P, _P, _U32, _P], ctypes.c_int),
    "mi_graph_create": ([_P, _U32, _U32, _U32, ctypes.POINTER(_P)], ctypes.c_int),
    "mi_graph_run": ([_P, _P, _U32, _P, _P, _U32, _U32, _P, ctypes.POINTER(MiStats)],
                     ctypes.c_int),
    "mi_graph_destroy": ([_P], ctypes.c_int),
    "mi_graph_launch": ([_P, _P, _U32, _P, _P, _U32, _U32], ctypes.c_int),
    "mi_graph_finish": ([_P, _P, ctypes.POINTER(MiStats)], ctypes.c_int),
    "mi_copy_calibration": ([_P, ctypes.c_uint64, _U32, ctypes.POINTER(ctypes.c_float)],
                            ctypes.c_int),
    "mi_valu_rate": ([_P, _U32, _U32, ctypes.POINTER(ctypes.c_double),
                      ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "mi_mlp_create": ([_P, _U32, _U32, ctypes.POINTER(_P)], ctypes.c_int),
    "mi_mlp_destroy": ([_P], ctypes.c_int),
    "mi_mlp_get_params": ([_P, _P], ctypes.c_int),
    "mi_mlp_set_params": ([_P, _P], ctypes.c_int),
    "mi_mlp_forward": ([_P, _P, _U32, _P], ctypes.c_int),
    "mi_mlp_gradients": ([_P, _P, _P, _U32, _P, _P], ctypes.c_int),
    "mi_mlp_train_step": ([_P, _P, _P, _U32, ctypes.c_float, _P], ctypes.c_int),
    "mi_mlp_adam_state": ([_P, _P, _P, _P], ctypes.c_int),
    "mi_mlp_set_dataset": ([_P, _P, _P, _U32], ctypes.c_int),
    "mi_mlp_train_epoch": ([_P, _U32, ctypes.c_float, _U32, _U32, _P, _P], ctypes.c_int),
    "mi_mlp_save": ([_P, ctypes.c_char_p], ctypes.c_int),
    "mi_mlp_load": ([_P, ctypes.c_char_p], ctypes.c_int),
}

MLP_PARAMS = 99073

_lib = None


def library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads the device library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python -m gpu_ai_amd.build`")
        lib = ctypes.CDLL(path)
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def as_positions(pos) -> np.ndarray:
    arr = np.ascontiguousarray(pos)
    if arr.dtype != POS_DTYPE:
        raise TypeError(f"positions must have dtype POS_DTYPE, got {arr.dtype}")
    return arr


@dataclass
class PlayoutResult:
    winners: np.ndarray
    stats: dict
    plies: np.ndarray | None = None
    hashes: np.ndarray | None = None


# Objects still open when the interpreter exits are closed by an atexit hook, children (graphs,
# networks) before their device contexts, while the HIP runtime is certainly alive: finalizers
# at interpreter shutdown run in no guaranteed order, possibly after native teardown.
_open_children: "weakref.WeakSet" = weakref.WeakSet()
_open_engines: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _close_open_objects() -> None:
    for group in (_open_children, _open_engines):
        for obj in list(group):
            try:
                obj.close()
            except Exception:
                pass


class Engine:
    """One device context (stream, events, buffers) of the playout engine."""

    def __init__(self, device: int = 0, block_size: int = 0, blocks_per_cu: int = 0,
                 refill: int = REFILL_QUEUE, defer_min: int = 0):
        self.lib = library()
        self.ctx = ctypes.c_void_p()
        t = MiTuning(block_size, blocks_per_cu, refill, defer_min)
        rc = self.lib.mi_create(ctypes.byref(self.ctx), device, ctypes.byref(t))
        if rc != MI_OK:
            raise MiError(rc, self.lib.mi_strerror(rc).decode())
        _open_engines.add(self)

    # -- helpers --
    def _check(self, rc: int) -> None:
        if rc != MI_OK:
            msg = self.lib.mi_last_error(self.ctx).decode() or self.lib.mi_strerror(rc).decode()
            raise MiError(rc, msg)

    def close(self) -> None:
        if self.ctx:
            self.lib.mi_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tuning(self) -> dict:
        t = MiTuning()
        self._check(self.lib.mi_get_tuning(self.ctx, ctypes.byref(t)))
        return {"block_size": t.block_size, "blocks_per_cu": t.blocks_per_cu, "refill": t.refill,
                "defer_min": t.defer_min}

    def set_tuning(self, block_size: int = 0, blocks_per_cu: int = 0, refill: int = REFILL_QUEUE,
                   defer_min: int = 0):
        t = MiTuning(block_size, blocks_per_cu, refill, defer_min)
        self._check(self.lib.mi_set_tuning(self.ctx, ctypes.byref(t)))

    def device_info(self) -> dict:
        cu, ws, clk = _U32(), _U32(), _U32()
        self._check(self.lib.mi_device_info(self.ctx, ctypes.byref(cu), ctypes.byref(ws),
                                            ctypes.byref(clk)))
        return {"compute_units": cu.value, "wave_size": ws.value, "clock_khz": clk.value}

    @staticmethod
    def _counts(npos: int, per_pos: int | None, counts) -> tuple[np.ndarray | None, int, int]:
        if counts is not None:
            c = np.ascontiguousarray(counts, dtype=np.uint32)
            if c.shape != (npos,):
                raise ValueError("counts must have one entry per position")
            return c, 0, int(c.sum(dtype=np.uint64))
        return None, int(per_pos), npos * int(per_pos)

    # -- entry points --
    def run_playouts(self, pos, per_pos: int | None = None, counts=None, seed: int = 12345,
                     epoch: int = 0, policy: int = POLICY_RANDOM, keys=None) -> PlayoutResult:
        """mi_run_playouts, or mi_run_playouts_keyed when `keys` (one u32 per position) is
        given."""
        pos = as_positions(pos)
        c, pp, total = self._counts(len(pos), per_pos, counts)
        winners = np.empty(total, dtype=np.int8)
        st = MiStats()
        if keys is None:
            self._check(self.lib.mi_run_playouts(self.ctx, _ptr(pos), len(pos), _ptr(c), pp, seed,
                                                 epoch, policy, _ptr(winners), ctypes.byref(st)))
        else:
            k = np.ascontiguousarray(keys, dtype=np.uint32)
            if k.shape != (len(pos),):
                raise ValueError("keys must have one entry per position")
            self._check(self.lib.mi_run_playouts_keyed(self.ctx, _ptr(pos), len(pos), _ptr(c), pp,
                                                       _ptr(k), seed, epoch, policy,
                                                       _ptr(winners), ctypes.byref(st)))
        return PlayoutResult(winners, st.as_dict())

    def run_traced(self, pos, per_pos: int | None = None, counts=None, seed: int = 12345,
                   epoch: int = 0, policy: int = POLICY_RANDOM) -> PlayoutResult:
        pos = as_positions(pos)
        c, pp, total = self._counts(len(pos), per_pos, counts)
        winners = np.empty(total, dtype=np.int8)
        plies = np.empty(total, dtype=np.uint32)
        hashes = np.empty(total, dtype=np.uint32)
        st = MiStats()
        self._check(self.lib.mi_run_playouts_traced(self.ctx, _ptr(pos), len(pos), _ptr(c), pp,
                                                    seed, epoch, policy, _ptr(winners),
                                                    _ptr(plies), _ptr(hashes), ctypes.byref(st)))
        return PlayoutResult(winners, st.as_dict(), plies, hashes)

    def run_tally(self, pos, per_pos: int | None = None, counts=None, seed: int = 12345,
                  epoch: int = 0, policy: int = POLICY_RANDOM) -> tuple[np.ndarray, dict]:
        pos = as_positions(pos)
        c, pp, _ = self._counts(len(pos), per_pos, counts)
        tallies = np.zeros((len(pos), 3), dtype=np.uint32)
        st = MiStats()
        self._check(self.lib.mi_run_playouts_tally(self.ctx, _ptr(pos), len(pos), _ptr(c), pp,
                                                   seed, epoch, policy, _ptr(tallies),
                                                   ctypes.byref(st)))
        return tallies, st.as_dict()

    def run_playouts_noalloc(self, d_pos: int, npos: int, d_counts: int | None, per_pos: int,
                             d_winners: int, seed: int = 12345, epoch: int = 0,
                             policy: int = POLICY_RANDOM, d_status: int | None = None,
                             stream: int | None = None) -> None:
        """mi_run_playouts_noalloc: device addresses in, winners left in device memory; enqueued
        on `stream` (a hipStream_t handle; None: the context's stream) without waiting."""
        self._check(self.lib.mi_run_playouts_noalloc(self.ctx, d_pos, npos, d_counts, per_pos,
                                                     seed, epoch, policy, d_winners, d_status,
                                                     stream))

    def gen_moves(self, pos, cap: int = 128) -> tuple[np.ndarray, np.ndarray]:
        pos = as_positions(pos)
        moves = np.zeros((len(pos), cap), dtype=MOVE_DTYPE)
        counts = np.zeros(len(pos), dtype=np.uint32)
        self._check(self.lib.mi_gen_moves(self.ctx, _ptr(pos), len(pos), _ptr(moves), cap,
                                          _ptr(counts)))
        return moves, counts

    def start_positions(self, seed: int, n: int, kmin: int, kmax: int,
                        reject_terminal: bool = False) -> np.ndarray:
        out = np.zeros(n, dtype=POS_DTYPE)
        self._check(self.lib.mi_gen_start_positions(self.ctx, seed, n, kmin, kmax,
                                                    int(reject_terminal), _ptr(out)))
        return out

    def stage(self, pos, per_pos: int | None = None, counts=None) -> int:
        pos = as_positions(pos)
        c, pp, total = self._counts(len(pos), per_pos, counts)
        self._check(self.lib.mi_stage(self.ctx, _ptr(pos), len(pos), _ptr(c), pp))
        self._staged_total = total
        return total

    def launch_staged(self, seed: int = 12345, epoch: int = 0,
                      policy: int = POLICY_RANDOM) -> dict:
        st = MiStats()
        self._check(self.lib.mi_launch_staged(self.ctx, seed, epoch, policy, ctypes.byref(st)))
        return st.as_dict()

    def fetch_winners(self) -> np.ndarray:
        w = np.empty(self._staged_total, dtype=np.int8)
        self._check(self.lib.mi_fetch_winners(self.ctx, _ptr(w), self._staged_total))
        return w

    def autotune(self, pos, per_pos: int | None = None, counts=None, policy: int = POLICY_RANDOM,
                 reps: int = 3) -> tuple[dict, list[dict]]:
        """Sweeps launch shapes on this batch; the fastest becomes the context's tuning."""
        pos = as_positions(pos)
        c, pp, _ = self._counts(len(pos), per_pos, counts)
        res = (MiTuneResult * 256)()
        nres = _U32()
        best = MiTuning()
        self._check(self.lib.mi_autotune(self.ctx, _ptr(pos), len(pos), _ptr(c), pp, policy, reps,
                                         res, 256, ctypes.byref(nres), ctypes.byref(best)))
        rows = [{"block_size": r.tuning.block_size, "blocks_per_cu": r.tuning.blocks_per_cu,
                 "refill": r.tuning.refill, "defer_min": r.tuning.defer_min,
                 "kernel_ms": r.kernel_ms, "playouts_per_s": r.playouts_per_s}
                for r in res[:min(nres.value, 256)]]
        return ({"block_size": best.block_size, "blocks_per_cu": best.blocks_per_cu,
                 "refill": best.refill, "defer_min": best.defer_min}, rows)

    def autotune_sizes(self, pos, sizes, policy: int = POLICY_RANDOM, reps: int = 3) -> list[dict]:
        """Fills the launch-shape table for these batch sizes (the first min(size, len(pos))
        positions, equal playout counts); returns the fastest shape of each size."""
        pos = as_positions(pos)
        sz = np.ascontiguousarray(sizes, dtype=np.uint32)
        best = (MiTuneEntry * len(sz))()
        self._check(self.lib.mi_autotune_sizes(self.ctx, _ptr(pos), len(pos), policy, _ptr(sz),
                                               len(sz), reps, best))
        return [b.as_dict() for b in best]

    def tuning_table(self) -> list[dict]:
        """The launch-shape table's entries (policy, bucket, refill order)."""
        n = _U32()
        self._check(self.lib.mi_get_tuning_table(self.ctx, None, 0, ctypes.byref(n)))
        out = (MiTuneEntry * max(n.value, 1))()
        self._check(self.lib.mi_get_tuning_table(self.ctx, out, n.value, ctypes.byref(n)))
        return [e.as_dict() for e in out[:n.value]]

    def graph(self, max_positions: int, max_playouts: int,
              policy: int = POLICY_RANDOM) -> "GraphBatch":
        return GraphBatch(self, max_positions, max_playouts, policy)

    def copy_calibration(self, nbytes: int, repeats: int = 5) -> float:
        ms = ctypes.c_float()
        self._check(self.lib.mi_copy_calibration(self.ctx, nbytes, repeats, ctypes.byref(ms)))
        return ms.value

    def valu_rate(self, iters: int = 16384, repeats: int = 5) -> tuple[float, float]:
        """Measured int32 VALU issue rate (lane-ops/s, best of `repeats`) and average ms."""
        rate, ms = ctypes.c_double(), ctypes.c_float()
        self._check(self.lib.mi_valu_rate(self.ctx, iters, repeats, ctypes.byref(rate),
                                          ctypes.byref(ms)))
        return rate.value, ms.value

    def check_divmod(self, r, n) -> tuple[np.ndarray, np.ndarray]:
        r = np.ascontiguousarray(r, dtype=np.uint32)
        n = np.ascontiguousarray(n, dtype=np.uint32)
        q = np.empty_like(r)
        rem = np.empty_like(r)
        self._check(self.lib.mi_check_divmod(self.ctx, _ptr(r), _ptr(n), len(r), _ptr(q),
                                             _ptr(rem)))
        return q, rem


class GraphBatch:
    """A hipGraph of (reset, persistent playout kernel, copy out) replayed per batch, after an
    asynchronous copy of the batch; the playout node's grid is sized to each batch."""

    def __init__(self, engine: Engine, max_positions: int, max_playouts: int, policy: int):
        self.engine = engine
        self.handle = ctypes.c_void_p()
        engine._check(engine.lib.mi_graph_create(engine.ctx, max_positions, max_playouts, policy,
                                                 ctypes.byref(self.handle)))
        _open_children.add(self)

    @staticmethod
    def _keys(keys, npos: int):
        if keys is None:
            return None
        k = np.ascontiguousarray(keys, dtype=np.uint32)
        if k.shape != (npos,):
            raise ValueError("keys must have one entry per position")
        return k

    def run(self, pos, counts, seed: int = 12345, epoch: int = 0, keys=None) -> PlayoutResult:
        pos = as_positions(pos)
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        k = self._keys(keys, len(pos))
        winners = np.empty(int(c.sum(dtype=np.uint64)), dtype=np.int8)
        st = MiStats()
        self.engine._check(self.engine.lib.mi_graph_run(self.handle, _ptr(pos), len(pos), _ptr(c),
                                                        _ptr(k), seed, epoch, _ptr(winners),
                                                        ctypes.byref(st)))
        return PlayoutResult(winners, st.as_dict())

    def launch(self, pos, counts, seed: int = 12345, epoch: int = 0, keys=None) -> None:
        """Enqueues one replay and returns at once (the batch is copied before returning)."""
        pos = as_positions(pos)
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        k = self._keys(keys, len(pos))
        self._total = int(c.sum(dtype=np.uint64))
        self.engine._check(self.engine.lib.mi_graph_launch(self.handle, _ptr(pos), len(pos),
                                                           _ptr(c), _ptr(k), seed, epoch))

    def finish(self) -> PlayoutResult:
        """Waits for the replay enqueued by launch and returns its winners."""
        winners = np.empty(getattr(self, "_total", 0), dtype=np.int8)
        st = MiStats()
        self.engine._check(self.engine.lib.mi_graph_finish(self.handle, _ptr(winners),
                                                           ctypes.byref(st)))
        return PlayoutResult(winners, st.as_dict())

    def close(self) -> None:
        if self.handle:
            self.engine.lib.mi_graph_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class ValueNet:
    """The 128-256-256-1 value network of one device context (include/gpuai/mi_mlp.h)."""

    def __init__(self, engine: Engine, max_batch: int = 4096, seed: int = 1):
        self.engine = engine
        self.lib = engine.lib
        self.handle = ctypes.c_void_p()
        engine._check(self.lib.mi_mlp_create(engine.ctx, max_batch, seed,
                                             ctypes.byref(self.handle)))
        self.max_batch = max_batch
        _open_children.add(self)

    def _check(self, rc: int) -> None:
        self.engine._check(rc)

    def close(self) -> None:
        if self.handle:
            self.lib.mi_mlp_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        w = np.empty(MLP_PARAMS, dtype=np.float32)
        self._check(self.lib.mi_mlp_get_params(self.handle, _ptr(w)))
        return w

    def set_params(self, w) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.shape != (MLP_PARAMS,):
            raise ValueError(f"expected {MLP_PARAMS} parameters")
        self._check(self.lib.mi_mlp_set_params(self.handle, _ptr(w)))

    def forward(self, pos) -> np.ndarray:
        pos = as_positions(pos)
        out = np.empty(len(pos), dtype=np.float32)
        self._check(self.lib.mi_mlp_forward(self.handle, _ptr(pos), len(pos), _ptr(out)))
        return out

    def gradients(self, pos, targets) -> tuple[np.ndarray, float]:
        pos = as_positions(pos)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self.lib.mi_mlp_gradients(self.handle, _ptr(pos), _ptr(t), len(pos), _ptr(g),
                                              ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, pos, targets, lr: float = 1e-3) -> float:
        pos = as_positions(pos)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self.lib.mi_mlp_train_step(self.handle, _ptr(pos), _ptr(t), len(pos), lr,
                                               ctypes.byref(loss)))
        return loss.value

    def adam_state(self) -> tuple[np.ndarray, np.ndarray, int]:
        m = np.empty(MLP_PARAMS, dtype=np.float32)
        v = np.empty(MLP_PARAMS, dtype=np.float32)
        step = _U32()
        self._check(self.lib.mi_mlp_adam_state(self.handle, _ptr(m), _ptr(v), ctypes.byref(step)))
        return m, v, step.value

    def set_dataset(self, pos, targets) -> None:
        pos = as_positions(pos)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        if t.shape != (len(pos),):
            raise ValueError("one target per position")
        self._check(self.lib.mi_mlp_set_dataset(self.handle, _ptr(pos), _ptr(t), len(pos)))

    def train_epoch(self, batch: int = 4096, lr: float = 1e-3, shuffle_seed: int = 7,
                    epoch: int = 0) -> tuple[float, float]:
        loss, ms = ctypes.c_float(), ctypes.c_float()
        self._check(self.lib.mi_mlp_train_epoch(self.handle, batch, lr, shuffle_seed, epoch,
                                                ctypes.byref(loss), ctypes.byref(ms)))
        return loss.value, ms.value

    def save(self, path: str) -> None:
        self._check(self.lib.mi_mlp_save(self.handle, os.fsencode(path)))

    def load(self, path: str) -> None:
        self._check(self.lib.mi_mlp_load(self.handle, os.fsencode(path)))
