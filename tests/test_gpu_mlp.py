"""Device value network (include/gpuai/mi_mlp.h) against its CPU restatement (oracle/mlp_ref.py).

Tolerances (floating point: BF16 operands, FP32 accumulation in a different order than the
oracle's float64 sums, so an occasional stored activation differs by one BF16 step):
  forward values      max |dy| <= 1e-3, mean |dy| <= 1e-5
  gradients           per tensor: max |dg| <= 2e-3 * max |g|, cosine >= 0.99999; loss rel 1e-4
  Adam update         given the device's own gradient: |dw| <= 1e-6, moments rel 1e-5
                      (absolute slack 1e-7 of the largest moment)
Integer-exact properties (initialisation, parameter round trips, colour symmetry, epoch vs
steps, checkpoints) are checked bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu

TENSORS = {"W1": (M.OFF_W1, M.OFF_B1), "b1": (M.OFF_B1, M.OFF_W2), "W2": (M.OFF_W2, M.OFF_B2),
           "b2": (M.OFF_B2, M.OFF_W3), "w3": (M.OFF_W3, M.OFF_B3), "b3": (M.OFF_B3, M.N_PARAMS)}


def _params(seed=5):
    rng = np.random.default_rng(seed)
    w = np.zeros(M.N_PARAMS, np.float32)
    w[M.OFF_W1:M.OFF_B1] = rng.uniform(-0.2, 0.2, 32768)
    w[M.OFF_B1:M.OFF_W2] = rng.uniform(-0.1, 0.1, 256)
    w[M.OFF_W2:M.OFF_B2] = rng.uniform(-0.15, 0.15, 65536)
    w[M.OFF_B2:M.OFF_W3] = rng.uniform(-0.1, 0.1, 256)
    w[M.OFF_W3:M.OFF_B3] = rng.uniform(-0.15, 0.15, 256)
    w[M.OFF_B3] = 0.05
    return w


@pytest.fixture(scope="module")
def positions(oracle):
    # openings to endgames (kings appear late), non-terminal
    a = oracle.start_positions(21, 6200, 0, 40, reject_terminal=True)
    b = oracle.start_positions(22, 6200, 40, 120, reject_terminal=True)
    rng = np.random.default_rng(23)
    return np.concatenate([a, b])[rng.permutation(12400)]  # every batch mixes both


def _net(engine, max_batch=1024, seed=1):
    from gpu_ai_amd.engine import ValueNet
    return ValueNet(engine, max_batch, seed)


def test_mlp_init_matches_oracle(engine, oracle):
    with _net(engine, 256, seed=42) as net:
        np.testing.assert_array_equal(net.get_params(), M.init_params(42, oracle.philox))
        m, v, step = net.adam_state()
        assert step == 0 and not m.any() and not v.any()


def test_mlp_params_round_trip(engine):
    w = _params(3)
    with _net(engine, 64) as net:
        net.set_params(w)
        np.testing.assert_array_equal(net.get_params(), w)


@pytest.mark.parametrize("n", [1, 64, 1000, 3000, 4096])
def test_mlp_forward_matches_oracle(engine, positions, n):
    w = _params(5)
    pos = positions[:n]
    with _net(engine, 4096) as net:
        net.set_params(w)
        y = net.forward(pos)
    ref = M.forward(w, M.encode(pos))["y"]
    d = np.abs(y - ref)
    print(f"n={n} forward max {d.max():.3g} mean {d.mean():.3g}")
    assert d.max() <= 1e-3 and d.mean() <= 1e-5


def test_mlp_forward_empty_and_oversized(engine, positions):
    from gpu_ai_amd.engine import MiError
    with _net(engine, 64) as net:
        assert net.forward(positions[:0]).shape == (0,)
        with pytest.raises(MiError) as e:
            net.forward(positions[:65])
        assert e.value.code == 1


def test_mlp_forward_colour_symmetry(engine, positions):
    """A position and its colour-exchanged, rotated twin have the same value for the mover."""
    pos = positions[:512]
    with _net(engine, 512, seed=9) as net:
        np.testing.assert_array_equal(net.forward(pos), net.forward(M.mirror(pos)))


@pytest.mark.parametrize("n", [200, 300, 512, 2048, 4096])  # 300: a half-filled last split
def test_mlp_gradients_match_oracle(engine, positions, n):
    w = _params(6)
    pos = positions[:n]
    t = np.random.default_rng(n).uniform(-1, 1, n).astype(np.float32)
    with _net(engine, 4096) as net:
        net.set_params(w)
        g, loss = net.gradients(pos, t)
        np.testing.assert_array_equal(net.get_params(), w)  # nothing updated
    ref, ref_loss = M.gradients(w, M.encode(pos), t)
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    for name, (lo, hi) in TENSORS.items():
        a, b = g[lo:hi].astype(np.float64), ref[lo:hi]
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
        cos = np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b))
        print(f"n={n} {name}: max rel {err:.3g} cosine {cos:.8f}")
        assert err <= 2e-3 and cos >= 0.99999, name


@pytest.mark.parametrize("batch", [512, 4096])
def test_mlp_adam_steps_match_oracle(engine, positions, batch):
    w = _params(7)
    opt = M.Adam(M.N_PARAMS, lr=1e-3)
    with _net(engine, batch) as net:
        net.set_params(w)
        ref_w = w.copy()
        for k in range(3):
            pos = positions[k * batch:(k + 1) * batch]
            t = np.random.default_rng(k).uniform(-1, 1, batch).astype(np.float32)
            g, _ = net.gradients(pos, t)
            net.train_step(pos, t, 1e-3)
            ref_w = opt.step(ref_w, g)
            got = net.get_params()
            np.testing.assert_allclose(got, ref_w, rtol=0, atol=1e-6)
            m, v, step = net.adam_state()
            assert step == k + 1
            # the device may fuse beta * m + (1 - beta) * g into one FMA: absolute slack at 1e-7
            # of the largest moment covers the cancellation in near-zero moments
            np.testing.assert_allclose(m, opt.m, rtol=1e-5, atol=1e-7 * np.abs(opt.m).max())
            np.testing.assert_allclose(v, opt.v, rtol=1e-5, atol=1e-7 * np.abs(opt.v).max())
            ref_w = got  # continue from the device's parameters


@pytest.mark.parametrize("batch,n", [(256, 1100), (4096, 12400)])
def test_mlp_epoch_equals_steps(engine, oracle, positions, batch, n):
    pos = positions[:n]
    t = np.random.default_rng(11).uniform(-1, 1, len(pos)).astype(np.float32)
    with _net(engine, batch, seed=4) as a, _net(engine, batch, seed=4) as b:
        a.set_dataset(pos, t)
        mean_loss, ms = a.train_epoch(batch, 1e-3, shuffle_seed=5, epoch=2)
        order = M.epoch_order(len(pos), 5, 2, oracle.philox)
        losses = []
        for k in range(len(pos) // batch):  # the trailing partial batch is dropped
            idx = order[k * batch:(k + 1) * batch]
            losses.append(b.train_step(pos[idx], t[idx], 1e-3))
        np.testing.assert_array_equal(a.get_params(), b.get_params())
        assert abs(mean_loss - np.mean(losses)) <= 1e-5 * np.mean(losses)
        assert ms > 0


def test_mlp_training_reduces_loss(engine, oracle):
    """Fitting playout win fractions: the held-out error falls well below the constant
    predictor's."""
    pos = oracle.start_positions(31, 12288, 0, 80, reject_terminal=True)
    tallies, _ = engine.run_tally(pos, per_pos=64, seed=77)
    mover = pos["turn"].astype(int)
    wins = np.where(mover == 0, tallies[:, 1], tallies[:, 2]).astype(np.float32)
    losses = np.where(mover == 0, tallies[:, 2], tallies[:, 1]).astype(np.float32)
    t = (wins - losses) / 64.0
    tr, te = slice(0, 10240), slice(10240, None)
    with _net(engine, 1024, seed=2) as net:
        net.set_dataset(pos[tr], t[tr])
        first = None
        for ep in range(15):
            loss, _ = net.train_epoch(512, 1e-3, shuffle_seed=1, epoch=ep)
            first = loss if first is None else first
        held = pos[te]
        y = np.concatenate([net.forward(held[a:a + 1024]) for a in range(0, len(held), 1024)])
    mse = float(np.mean((y - t[te]) ** 2))
    const = float(np.mean((t[tr].mean() - t[te]) ** 2))
    # what a leaf gets without the network: a 1- or 4-playout estimate (independent streams)
    est = {}
    for k in (1, 4):
        tk, _ = engine.run_tally(pos[te], per_pos=k, seed=500 + k)
        wk = np.where(mover[te] == 0, tk[:, 1], tk[:, 2]).astype(np.float32)
        lk = np.where(mover[te] == 0, tk[:, 2], tk[:, 1]).astype(np.float32)
        est[k] = float(np.mean(((wk - lk) / k - t[te]) ** 2))
    print(f"first epoch {first:.4f} last {loss:.4f} held-out {mse:.4f} constant {const:.4f} "
          f"1 playout {est[1]:.4f} 4 playouts {est[4]:.4f}")
    assert loss < 0.8 * first
    assert mse < 0.8 * const
    assert mse < est[4] < est[1]


def test_mlp_checkpoint_round_trip(engine, positions, tmp_path):
    pos = positions[:256]
    t = np.random.default_rng(0).uniform(-1, 1, 256).astype(np.float32)
    path = str(tmp_path / "value.ckpt")
    with _net(engine, 256, seed=8) as a, _net(engine, 256, seed=9) as b:
        a.train_step(pos, t)
        a.train_step(pos, t)
        a.save(path)
        b.load(path)
        np.testing.assert_array_equal(a.get_params(), b.get_params())
        ma, va, sa = a.adam_state()
        mb, vb, sb = b.adam_state()
        np.testing.assert_array_equal(ma, mb)
        np.testing.assert_array_equal(va, vb)
        assert sa == sb == 2
        np.testing.assert_array_equal(a.forward(pos), b.forward(pos))
        # the next step continues identically
        a.train_step(pos, t)
        b.train_step(pos, t)
        np.testing.assert_array_equal(a.get_params(), b.get_params())


def test_mlp_epoch_rejects_bad_batch(engine, positions):
    from gpu_ai_amd.engine import MiError
    with _net(engine, 256) as net:
        net.set_dataset(positions[:300], np.zeros(300, np.float32))
        for bad in (100, 512):
            with pytest.raises(MiError):
                net.train_epoch(bad)


def test_net_match_small(engine):
    """The match harness plays every game to its end and accounts for every leaf."""
    from gpu_ai_amd import host_api
    w = M.init_params(3, _philox())
    r = host_api.net_match(w, games=8, seed=5, opening_min=2, opening_max=6, move_ms=2.0,
                           leaf_batch=256, max_plies=60, net_batch=256)
    print({k: v for k, v in r.items() if k != "outcome"})
    assert r["games"] == 8
    assert r["net_wins"] + r["playout_wins"] + r["draws"] == 8
    o = r["outcome"]
    assert (o == 1).sum() == r["net_wins"] and (o == -1).sum() == r["playout_wins"]
    assert r["net_leaves"] > 0 and r["playout_leaves"] > 0
    assert r["net_iterations"] > 0 and r["playout_iterations"] > 0
    assert abs(r["score"] - (r["net_wins"] + 0.5 * r["draws"]) / 8) < 1e-12


def _philox():
    from oracle.pyoracle import Oracle
    return Oracle().philox
