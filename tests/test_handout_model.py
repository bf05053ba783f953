"""Host model of k_playouts' queue hand-out (gpu_ai_amd/csrc/playout_kernels.hip: slab_index and
the refill step of k_playouts; the slab layout is described in gpu_ai_amd/csrc/kernels.h), run
under random interleavings of the waves' claims.

Each wave is a generator that yields after its atomic claim on its slab's counter and after every
refill round, so a random scheduler explores orders in which waves of one slab overtake one
another.  The checks: every queue index below `total` starts exactly once, nothing past it starts,
and every wave ends.
"""
from __future__ import annotations

import random

import pytest

CHUNK = 16
SLAB_BLOCK = 128
SLABS = 16


def slab_index(c: int, s: int, ns: int) -> int:
    return ((c // SLAB_BLOCK) * ns + s) * SLAB_BLOCK | (c % SLAB_BLOCK)


def grid_size(need: int, resident: int) -> int:
    g = min(need, resident)
    if g > SLABS:
        g = -(-g // SLABS) * SLABS
    return max(g, 1)


class Shared:
    def __init__(self, nslabs: int):
        self.ctr = [0] * nslabs
        self.atomics = 0


def wave(sh: Shared, rng: random.Random, slab: int, ns: int, total: int, started: list,
         lanes: int = 64, mean_len: int = 3):
    remaining = [0] * lanes  # refill rounds left in each lane's playout; 0 = idle
    res_next = res_end = 0
    exhausted = False
    while True:
        if not exhausted:
            idle = [i for i in range(lanes) if remaining[i] == 0]
            if idle:
                cnt, avail = len(idle), res_end - res_next
                c = want = 0
                if avail < cnt:
                    want = -(-(cnt - avail) // CHUNK) * CHUNK
                    c = sh.ctr[slab]  # atomicAdd
                    sh.ctr[slab] += want
                    sh.atomics += 1
                    yield
                for rank, ln in enumerate(idle):
                    local = res_next + rank if rank < avail else c + (rank - avail)
                    idx = slab_index(local, slab, ns)
                    if idx < total:
                        started.append(idx)
                        remaining[ln] = 1 + rng.randrange(2 * mean_len)
                if avail < cnt:
                    res_next, res_end = c + (cnt - avail), c + want
                else:
                    res_next += cnt
                if slab_index(res_next, slab, ns) >= total:
                    exhausted = True
        if not any(remaining):
            return
        for i in range(lanes):
            if remaining[i]:
                remaining[i] -= 1
        yield


def run_model(seed: int, total: int, grid: int, waves_per_group: int, lanes: int = 64):
    rng = random.Random(seed)
    ns = min(SLABS, grid)
    sh = Shared(ns)
    started: list[int] = []
    live = [wave(sh, rng, g % ns, ns, total, started, lanes)
            for g in range(grid) for _ in range(waves_per_group)]
    steps = 0
    while live:
        k = rng.randrange(len(live))
        try:
            next(live[k])
        except StopIteration:
            live.pop(k)
        steps += 1
        assert steps < 50_000_000, "model did not terminate"
    return started, sh


@pytest.mark.parametrize("seed,total,block,resident", [
    (1, 0, 64, 3072), (2, 1, 64, 3072), (3, 1000, 64, 3072), (4, 4096, 64, 3072),
    (5, 65536, 1024, 256), (6, 20000, 768, 256), (7, 5000, 256, 3), (8, 777, 1024, 1),
    (9, 131072, 64, 48), (10, 2176, 128, 40), (11, 300000, 512, 512)])
def test_every_index_starts_once(seed, total, block, resident):
    grid = grid_size(-(-max(total, 1) // block), resident)
    assert grid <= SLABS or grid % SLABS == 0
    started, _ = run_model(seed, total, grid, block // 64)
    assert len(started) == total
    assert sorted(started) == list(range(total))


def test_slab_blocks_partition_the_queue():
    """Every queue index is one slab's: the slabs' blocks of 128 interleave without gaps."""
    for ns in (1, 3, 16):
        seen = sorted(slab_index(c, s, ns) for s in range(ns) for c in range(1024))
        assert seen == list(range(1024 * ns))


def test_claims_per_wave():
    """One device atomic per claim of whole chunks: at most total / 16 plus one per wave."""
    total = 1 << 17
    grid = grid_size(total // 1024, 256)
    started, sh = run_model(12, total, grid, 16)
    assert sorted(started) == list(range(total))
    assert sh.atomics <= total // CHUNK + grid * 16
