"""mi_run_playouts_noalloc (device buffers in and out, scratch in a fixed device array) against
mi_run_playouts on the same batches: two calls in a row, the second with a smaller batch."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from gpu_ai_amd.engine import (MI_ERR_INVALID_ARGUMENT, POLICY_HEURISTIC, POLICY_HEURISTIC_NOISY,
                               POLICY_RANDOM, MiError)

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


def _noalloc(engine, pos, per_pos, counts, seed, epoch, policy):
    """One allocation-free call on a stream of its own; returns (winners, status)."""
    from devmem import DeviceArray, Stream
    total = int(counts.sum()) if counts is not None else len(pos) * per_pos
    d_pos = DeviceArray(pos)
    d_counts = DeviceArray(counts.astype(np.uint32)) if counts is not None else None
    d_win = DeviceArray(np.full(max(total, 1), 7, dtype=np.int8))
    d_status = DeviceArray(np.full(1, 99, dtype=np.uint32))
    try:
        with Stream() as s:
            engine.run_playouts_noalloc(d_pos.address, len(pos),
                                        d_counts.address if d_counts is not None else None,
                                        per_pos, d_win.address, seed=seed, epoch=epoch,
                                        policy=policy, d_status=d_status.address,
                                        stream=s.handle.value)
            s.synchronize()
        return d_win.download(np.int8, total), int(d_status.download(np.uint32, 1)[0])
    finally:
        for b in (d_pos, d_counts, d_win, d_status):
            if b is not None:
                b.free()


@pytest.mark.parametrize("policy", [POLICY_RANDOM, POLICY_HEURISTIC, POLICY_HEURISTIC_NOISY])
def test_noalloc_equals_run_playouts_twice(engine, golden, policy):
    pos = golden["positions"]["open1024" if policy == POLICY_RANDOM else "mid1024"]
    rng = np.random.default_rng(40 + policy)
    cases = [
        # large, uniform counts: longest-first hand-out from the scratch's sort
        ("1024 positions x 512", pos, 512, None, 91, 3),
        # smaller, uneven counts with zeros: offsets from the scratch
        ("300 positions, uneven", pos[:300], 0,
         np.where(rng.random(300) < 0.2, 0, rng.integers(1, 90, 300)).astype(np.uint32), 92, 4),
    ]
    for name, p, per, counts, seed, epoch in cases:
        announce(f"noalloc {name}, policy {policy}")
        got, status = _noalloc(engine, p, per, counts, seed, epoch, policy)
        assert status == 0
        ref = engine.run_playouts(p, per_pos=per if counts is None else None, counts=counts,
                                  seed=seed, epoch=epoch, policy=policy)
        assert np.array_equal(got, ref.winners), name


def test_noalloc_ordered_uneven_and_rejections(engine, golden):
    pos = np.concatenate([golden["positions"]["open1024"], golden["positions"]["mid1024"]])
    counts = np.random.default_rng(5).integers(0, 200, len(pos)).astype(np.uint32)
    announce(f"noalloc ordered uneven: {int(counts.sum())} playouts")
    got, status = _noalloc(engine, pos, 0, counts, 17, 2, POLICY_RANDOM)
    assert status == 0
    ref = engine.run_playouts(pos, counts=counts, seed=17, epoch=2)
    assert np.array_equal(got, ref.winners)
    announce("noalloc: malformed position reported in the status word, nothing played")
    bad = pos[:8].copy()
    bad["p2"][3] |= bad["p1"][3] & (~bad["p1"][3] + 1)  # a square held by both sides
    got, status = _noalloc(engine, bad, 4, None, 1, 0, POLICY_RANDOM)
    assert status == MI_ERR_INVALID_ARGUMENT and np.all(got == 7)
    announce("noalloc: more positions than the scratch holds")
    with pytest.raises(MiError) as e:
        engine.run_playouts_noalloc(1 << 20, 65537, None, 1, 1 << 20)
    assert e.value.code == MI_ERR_INVALID_ARGUMENT and "65536" in str(e.value)
