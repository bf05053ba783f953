"""ctypes bindings of the oracle (TEST INFRASTRUCTURE ONLY).

Wraps oracle/liboracle.so (the C restatement in oracle.c) and, where it was built,
oracle/_ref/libref.so / libref_assert.so (the reference's own host code plus ref_harness.cpp).
Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py import this module.
Array layouts are those of gpu_ai_amd.engine (POS_DTYPE, MOVE_DTYPE).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")
REF_ASSERT_SO = os.path.join(HERE, "_ref", "libref_assert.so")

POS_DTYPE = np.dtype([("p1", "<u4"), ("p2", "<u4"), ("kings", "<u4"), ("turn", "u1"),
                      ("counter", "u1"), ("pad", "u1", (2,))])
MOVE_DTYPE = np.dtype([("from", "u1"), ("to", "u1"), ("jumps", "u1"), ("promoted", "u1"),
                       ("inter", "u1", (8,)), ("removed", "u1", (8,))])
TRACE_DTYPE = np.dtype([("winner", "i1"), ("max_n", "u1"), ("pad", "<u2"), ("plies", "<u4"),
                        ("hash", "<u4"), ("captures", "<u4")])

POLICY_RANDOM = 0
POLICY_HEURISTIC = 1        # without noise, ties broken by the ply's word
POLICY_HEURISTIC_NOISY = 2  # the reference's N(0, 0.11) draw on each move's weight
POLICY_RANDOM_LIBC = 9      # reference harness only: glibc rand() % n (timing, not reproducible)

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class _Lib:
    """Common wrapper; `prefix` is "or" for the restatement and "ref" for the reference."""

    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is missing: run `make -C oracle` (and `make -C oracle ref`)")
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        L, p = self.lib, prefix
        self._gen = getattr(L, f"{p}_gen_moves")
        self._gen.argtypes = [_P, _P, ctypes.c_int]
        self._gen.restype = ctypes.c_int
        self._perft = getattr(L, f"{p}_perft")
        self._perft.argtypes = [_P, ctypes.c_int]
        self._perft.restype = ctypes.c_uint64
        self._run = getattr(L, f"{p}_run_playouts")
        self._run.argtypes = [_P, _U32, _P, _U32, _U32, _U32, ctypes.c_int, _P, _P, ctypes.c_int]
        self._run.restype = ctypes.c_int
        self._start = getattr(L, f"{p}_gen_start_positions")
        self._start.argtypes = [_U32, _U32, _U32, _U32, ctypes.c_int, _P]
        self._start.restype = ctypes.c_int
        self._over = getattr(L, f"{p}_is_game_over")
        self._over.argtypes = [_P]
        self._over.restype = ctypes.c_int
        self._winner = getattr(L, f"{p}_winner")
        self._winner.argtypes = [_P]
        self._winner.restype = ctypes.c_int
        self._score_state = getattr(L, f"{p}_score_state")
        self._score_state.argtypes = [_P, _P]
        self._score_state.restype = None

    def gen_moves(self, pos: np.ndarray) -> np.ndarray:
        p = np.ascontiguousarray(pos, dtype=POS_DTYPE).reshape(1)
        out = np.zeros(256, dtype=MOVE_DTYPE)
        n = self._gen(_ptr(p), _ptr(out), 256)
        if n < 0:
            raise RuntimeError("move list longer than 256")
        return out[:n].copy()

    def gen_moves_many(self, pos: np.ndarray, cap: int = 128) -> tuple[np.ndarray, np.ndarray]:
        pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
        moves = np.zeros((len(pos), cap), dtype=MOVE_DTYPE)
        counts = np.zeros(len(pos), dtype=np.uint32)
        buf = np.zeros(256, dtype=MOVE_DTYPE)
        for i in range(len(pos)):
            n = self._gen(_ptr(pos[i:i + 1]), _ptr(buf), 256)
            counts[i] = n
            moves[i, :min(n, cap)] = buf[:min(n, cap)]
        return moves, counts

    def perft(self, pos: np.ndarray, depth: int) -> int:
        p = np.ascontiguousarray(pos, dtype=POS_DTYPE).reshape(1)
        return int(self._perft(_ptr(p), depth))

    def is_game_over(self, pos) -> bool:
        p = np.ascontiguousarray(pos, dtype=POS_DTYPE).reshape(1)
        return bool(self._over(_ptr(p)))

    def winner(self, pos) -> int:
        p = np.ascontiguousarray(pos, dtype=POS_DTYPE).reshape(1)
        return int(self._winner(_ptr(p)))

    def score_state(self, pos) -> tuple[int, int]:
        p = np.ascontiguousarray(pos, dtype=POS_DTYPE).reshape(1)
        s = np.zeros(2, dtype=np.uint32)
        self._score_state(_ptr(p), _ptr(s))
        return int(s[0]), int(s[1])

    def run_playouts(self, pos, per_pos: int | None = None, counts=None, seed: int = 12345,
                     epoch: int = 0, policy: int = POLICY_RANDOM, threads: int = 0,
                     traces: bool = True, keys=None):
        pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
        c = None if counts is None else np.ascontiguousarray(counts, dtype=np.uint32)
        total = int(c.sum()) if c is not None else len(pos) * int(per_pos)
        winners = np.empty(total, dtype=np.int8)
        tr = np.empty(total, dtype=TRACE_DTYPE) if traces else None
        pp = 0 if per_pos is None else int(per_pos)
        if keys is not None:  # stream keys (restatement only)
            k = np.ascontiguousarray(keys, dtype=np.uint32)
            if k.shape != (len(pos),):
                raise ValueError("one key per position")
            fn = self.lib.or_run_playouts_keyed
            fn.argtypes = [_P, _U32, _P, _U32, _P, _U32, _U32, ctypes.c_int, _P, _P, ctypes.c_int]
            fn.restype = ctypes.c_int
            rc = fn(_ptr(pos), len(pos), _ptr(c), pp, _ptr(k), seed, epoch, policy, _ptr(winners),
                    _ptr(tr), threads)
        else:
            rc = self._run(_ptr(pos), len(pos), _ptr(c), pp, seed, epoch, policy, _ptr(winners),
                           _ptr(tr), threads)
        if rc != 0:
            raise RuntimeError(f"{self.prefix}_run_playouts failed")
        return winners, tr

    def start_positions(self, seed: int, n: int, kmin: int, kmax: int,
                        reject_terminal: bool = False) -> np.ndarray:
        out = np.zeros(n, dtype=POS_DTYPE)
        rc = self._start(seed, n, kmin, kmax, int(reject_terminal), _ptr(out))
        if rc != 0:
            raise RuntimeError("start-position generation failed")
        return out


class Oracle(_Lib):
    def __init__(self, path: str = ORACLE_SO):
        super().__init__(path, "or")
        L = self.lib
        L.or_philox4x32_10.argtypes = [_P, _P, _P]
        L.or_philox4x32_10.restype = None
        L.or_stream_word.argtypes = [_U32] * 6
        L.or_stream_word.restype = _U32
        L.or_apply.argtypes = [_P, _P]
        L.or_apply.restype = ctypes.c_int
        L.or_start_position.argtypes = []
        L.or_score_move.argtypes = [_P, _P, _P]
        L.or_score_move.restype = None
        for name in ("or_fmix32",):
            getattr(L, name).argtypes = [_U32]
            getattr(L, name).restype = _U32
        L.or_move_digest.argtypes = [_P]
        L.or_move_digest.restype = _U32
        L.or_list_digest.argtypes = [_P, ctypes.c_int]
        L.or_list_digest.restype = _U32
        L.or_heuristic_noise.argtypes = [_U32, _U32]
        L.or_heuristic_noise.restype = ctypes.c_float
        L.or_normal_pairs.argtypes = [_P, _P, _U32, _P]
        L.or_normal_pairs.restype = None

    def choice_counts(self, pos, policy: int, seed: int, trials: int) -> np.ndarray:
        """How often each move (reference index) is chosen in `trials` plies from pos."""
        fn = self.lib.or_choice_counts
        fn.argtypes = [_P, ctypes.c_int, _U32, _U32, _P, ctypes.c_int]
        fn.restype = ctypes.c_int
        p = np.ascontiguousarray(pos, dtype=POS_DTYPE).reshape(1)
        out = np.zeros(256, dtype=np.uint32)
        n = fn(_ptr(p), policy, seed, trials, _ptr(out), 256)
        if n < 0:
            raise RuntimeError("move list longer than 256")
        return out[:n].copy()

    def heuristic_noise(self, r: int, m: int) -> float:
        """The noisy heuristic's draw for move m (mover order) in a ply of word r."""
        return float(self.lib.or_heuristic_noise(r, m))

    def normal_pairs(self, a, b) -> np.ndarray:
        """Box-Muller pairs of the word arrays a, b: shape (n, 2), float32."""
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        z = np.zeros((len(a), 2), dtype=np.float32)
        self.lib.or_normal_pairs(_ptr(a), _ptr(b), len(a), _ptr(z))
        return z

    def philox(self, ctr, key) -> np.ndarray:
        c = np.ascontiguousarray(ctr, dtype=np.uint32)
        k = np.ascontiguousarray(key, dtype=np.uint32)
        out = np.zeros(4, dtype=np.uint32)
        self.lib.or_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
        return out

    def stream_word(self, seed, pos, ply, playout, tag, epoch) -> int:
        return int(self.lib.or_stream_word(seed, pos, ply, playout, tag, epoch))

    def apply(self, pos, move) -> np.ndarray:
        p = np.array(pos, dtype=POS_DTYPE).reshape(1).copy()
        m = np.array(move, dtype=MOVE_DTYPE).reshape(1)
        if self.lib.or_apply(_ptr(p), _ptr(m)) != 0:
            raise ValueError("invalid move for this position")
        return p[0]

    def score_move(self, pos, move) -> tuple[int, int]:
        p = np.ascontiguousarray(pos, dtype=POS_DTYPE).reshape(1)
        m = np.array(move, dtype=MOVE_DTYPE).reshape(1)
        s = np.zeros(2, dtype=np.int32)
        self.lib.or_score_move(_ptr(p), _ptr(m), _ptr(s))
        return int(s[0]), int(s[1])

    def list_digest(self, moves: np.ndarray) -> int:
        m = np.ascontiguousarray(moves, dtype=MOVE_DTYPE)
        return int(self.lib.or_list_digest(_ptr(m), len(m)))


class Reference(_Lib):
    """The reference's own host code (available where oracle/_ref was built)."""

    def __init__(self, path: str = REF_ASSERT_SO):
        super().__init__(path, "ref")
        L = self.lib
        L.ref_apply_index.argtypes = [_P, ctypes.c_int]
        L.ref_apply_index.restype = ctypes.c_int
        L.ref_score_move_index.argtypes = [_P, ctypes.c_int, _P, _P]
        L.ref_score_move_index.restype = ctypes.c_int

    def mcts_trace(self, root, iterations: int, initial: int, scale: float) -> np.ndarray:
        """ref_mcts_trace: the reference's GameTree driven by a fixed winner function."""
        fn = self.lib.ref_mcts_trace
        fn.argtypes = [_P, _U32, _U32, ctypes.c_double, _P, _U32]
        fn.restype = ctypes.c_int
        p = np.ascontiguousarray(root, dtype=POS_DTYPE).reshape(1)
        out = np.zeros(5 * iterations + 4, dtype=np.uint32)
        n = fn(_ptr(p), iterations, initial, scale, _ptr(out), len(out))
        if n < 0:
            raise RuntimeError("ref_mcts_trace failed")
        return out[:n]

    def host_heuristic_driver(self, pos, per_pos: int) -> np.ndarray:
        """The reference's HostHeuristicPlayoutDriver::runPlayouts on per_pos copies of each
        position (position-major), one thread, its own default-seeded noise stream."""
        fn = self.lib.ref_host_heuristic_driver
        fn.argtypes = [_P, _U32, _U32, _P]
        fn.restype = ctypes.c_int
        pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
        out = np.zeros(len(pos) * per_pos, dtype=np.int8)
        if fn(_ptr(pos), len(pos), per_pos, _ptr(out)) != 0:
            raise RuntimeError("ref_host_heuristic_driver failed")
        return out

    def noisy_choice_counts(self, pos, trials: int, seed: int) -> np.ndarray:
        """One noisy heuristic ply from pos, `trials` times, with the reference's weights and
        its std::normal_distribution noise: how often each move (reference index) won."""
        fn = self.lib.ref_noisy_choice_counts
        fn.argtypes = [_P, _U32, _U32, _P, ctypes.c_int]
        fn.restype = ctypes.c_int
        p = np.ascontiguousarray(pos, dtype=POS_DTYPE).reshape(1)
        out = np.zeros(256, dtype=np.uint32)
        n = fn(_ptr(p), trials, seed, _ptr(out), 256)
        if n < 0:
            raise RuntimeError("move list longer than 256")
        return out[:n].copy()

    def apply_index(self, pos, index: int) -> np.ndarray:
        p = np.array(pos, dtype=POS_DTYPE).reshape(1).copy()
        if self.lib.ref_apply_index(_ptr(p), index) != 0:
            raise ValueError("index out of range")
        return p[0]

    def score_move_index(self, pos, index: int) -> tuple[int, int, float]:
        p = np.ascontiguousarray(pos, dtype=POS_DTYPE).reshape(1)
        s = np.zeros(2, dtype=np.int32)
        w = np.zeros(1, dtype=np.float32)
        if self.lib.ref_score_move_index(_ptr(p), index, _ptr(s), _ptr(w)) != 0:
            raise ValueError("index out of range")
        return int(s[0]), int(s[1]), float(w[0])


def reference_available(path: str = REF_ASSERT_SO) -> bool:
    return os.path.exists(path)


def start_position() -> np.ndarray:
    p = np.zeros(1, dtype=POS_DTYPE)[0]
    p["p1"] = 0x00000FFF
    p["p2"] = 0xFFF00000
    return p
