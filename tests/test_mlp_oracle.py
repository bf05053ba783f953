"""Pins the value network's CPU restatement (oracle/mlp_ref.py) before it checks the device.

Each oracle function is checked against something that fixes it independently of its own
formulas: torch's bfloat16 conversion (rounding), torch autograd on the same network in float64
(gradients), central finite differences of the loss (gradients, without torch), torch.optim.Adam
(update rule), closed forms of the encoding (start position, colour symmetry, plane counts), and
planted mistakes that each of those checks must catch.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import mlp_ref as M
from oracle.pyoracle import POS_DTYPE, start_position


def _positions(oracle, n=96, seed=3):
    return oracle.start_positions(seed, n, 0, 60, reject_terminal=True)


def _params(seed=5, scale=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(M.N_PARAMS, np.float32)
    w[M.OFF_W1:M.OFF_B1] = rng.uniform(-0.2, 0.2, 32768) * scale
    w[M.OFF_B1:M.OFF_W2] = rng.uniform(-0.1, 0.1, 256) * scale
    w[M.OFF_W2:M.OFF_B2] = rng.uniform(-0.15, 0.15, 65536) * scale
    w[M.OFF_B2:M.OFF_W3] = rng.uniform(-0.1, 0.1, 256) * scale
    w[M.OFF_W3:M.OFF_B3] = rng.uniform(-0.15, 0.15, 256) * scale
    w[M.OFF_B3] = 0.05
    return w


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(100000).astype(np.float32),
                        rng.uniform(-1e-3, 1e-3, 1000).astype(np.float32),
                        # exact ties: low 16 bits 0x8000, both parities of bit 16
                        (np.arange(1000, dtype=np.uint32) << 16 | 0x3F000000 | 0x8000).view(np.float32),
                        np.array([0.0, -0.0, 1.0, -1.0, 65504.0], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(M.bf16(x), ref)


def test_encode_start_position_closed_form():
    s = start_position()
    X = M.encode(np.array([s]))[0]
    own_men, own_kings, opp_men, opp_kings = X.reshape(4, 32)
    np.testing.assert_array_equal(own_men, (np.arange(32) < 12).astype(np.float32))
    np.testing.assert_array_equal(opp_men, (np.arange(32) >= 20).astype(np.float32))
    assert own_kings.sum() == 0 and opp_kings.sum() == 0
    # player 2 to move in the start position sees the same board from its side
    s2 = s.copy()
    s2["turn"] = 1
    np.testing.assert_array_equal(M.encode(np.array([s2]))[0], X)


def test_encode_colour_symmetry_and_counts(oracle):
    pos = _positions(oracle, 200, seed=11)
    X = M.encode(pos)
    np.testing.assert_array_equal(M.encode(M.mirror(pos)), X)
    occ = np.array([bin(int(a) | int(b)).count("1") for a, b in zip(pos["p1"], pos["p2"])])
    np.testing.assert_array_equal(X.sum(1), occ)
    kings = np.array([bin(int(k) & (int(a) | int(b))).count("1")
                      for k, a, b in zip(pos["kings"], pos["p1"], pos["p2"])])
    np.testing.assert_array_equal(X[:, 32:64].sum(1) + X[:, 96:128].sum(1), kings)
    # the mover's pieces are in planes 0-1
    own = np.where(pos["turn"] == 0, pos["p1"], pos["p2"])
    np.testing.assert_array_equal(X[:, :64].sum(1), [bin(int(v)).count("1") for v in own])


def _torch_grads(w, X, t):
    W = torch.tensor(w.astype(np.float64), requires_grad=True)
    W1 = W[M.OFF_W1:M.OFF_B1].view(128, 256)
    b1 = W[M.OFF_B1:M.OFF_W2]
    W2 = W[M.OFF_W2:M.OFF_B2].view(256, 256)
    b2 = W[M.OFF_B2:M.OFF_W3]
    w3 = W[M.OFF_W3:M.OFF_B3]
    b3 = W[M.OFF_B3]
    x = torch.tensor(X.astype(np.float64))
    h1 = torch.relu(x @ W1 + b1)
    h2 = torch.relu(h1 @ W2 + b2)
    y = torch.tanh(h2 @ w3 + b3)
    loss = torch.mean((y - torch.tensor(t.astype(np.float64))) ** 2)
    loss.backward()
    return W.grad.numpy(), float(loss.detach())


def test_gradients_match_torch_autograd(oracle):
    pos = _positions(oracle, 96)
    X = M.encode(pos)
    t = np.random.default_rng(1).uniform(-1, 1, len(pos)).astype(np.float32)
    w = _params()
    g, loss = M.gradients(w.astype(np.float64), X, t, rounding=False)
    gt, lt = _torch_grads(w, X, t)
    assert abs(loss - lt) < 1e-12
    np.testing.assert_allclose(g, gt, rtol=1e-9, atol=1e-12)


def test_gradients_match_finite_differences(oracle):
    pos = _positions(oracle, 64, seed=9)
    X = M.encode(pos)
    t = np.random.default_rng(2).uniform(-1, 1, len(pos))
    w = _params(7).astype(np.float64)
    g, _ = M.gradients(w, X, t, rounding=False)
    rng = np.random.default_rng(3)
    idx = []
    for lo, hi in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
                   (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAMS)):
        cand = np.arange(lo, hi)
        big = cand[np.argsort(-np.abs(g[lo:hi]))[:4]]  # the coordinates that matter most
        idx += list(big) + list(rng.choice(cand, min(3, hi - lo), replace=False))
    def pattern(v):
        f = M.forward(v, X, rounding=False)
        return np.concatenate([(f["H1"] > 0).ravel(), (f["H2"] > 0).ravel()])

    base = pattern(w)
    for i in idx:
        # the loss is smooth while no ReLU changes state: shrink the step until none does
        h = 1e-6
        while True:
            wp, wm = w.copy(), w.copy()
            wp[i] += h
            wm[i] -= h
            if (np.array_equal(pattern(wp), base) and np.array_equal(pattern(wm), base)) or h < 1e-10:
                break
            h /= 4
        assert h >= 1e-10, f"coordinate {i} sits on a ReLU kink"
        fd = (M.loss_only(wp, X, t) - M.loss_only(wm, X, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, h, fd, g[i])


def test_rounded_gradients_close_to_exact(oracle):
    """BF16 operands perturb the gradient only at the level of BF16 resolution."""
    pos = _positions(oracle, 256, seed=4)
    X = M.encode(pos)
    t = np.random.default_rng(4).uniform(-1, 1, len(pos)).astype(np.float32)
    w = _params(8)
    g, _ = M.gradients(w, X, t, rounding=True)
    ge, _ = M.gradients(w.astype(np.float64), X, t, rounding=False)
    for lo, hi in ((M.OFF_W1, M.OFF_B1), (M.OFF_W2, M.OFF_B2), (M.OFF_W3, M.OFF_B3)):
        a, b = g[lo:hi], ge[lo:hi]
        assert np.abs(a - b).max() <= 0.1 * np.abs(b).max()
        assert np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.999


def test_adam_matches_torch():
    rng = np.random.default_rng(5)
    w0 = rng.standard_normal(1000).astype(np.float32)
    grads = [rng.standard_normal(1000).astype(np.float32) * s for s in (1.0, 0.1, 3.0, 1e-4)]
    opt = M.Adam(1000, lr=1e-3)
    w = w0.copy()
    p = torch.nn.Parameter(torch.from_numpy(w0.copy()))
    topt = torch.optim.Adam([p], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for g in grads:
        w = opt.step(w, g)
        p.grad = torch.from_numpy(g.copy())
        topt.step()
        np.testing.assert_allclose(w, p.detach().numpy(), rtol=0, atol=2e-7)


def test_planted_mistakes_are_caught(oracle):
    """A transposed W2 in the backward pass or a dropped tanh derivative fails the torch check."""
    pos = _positions(oracle, 64, seed=12)
    X = M.encode(pos)
    t = np.random.default_rng(6).uniform(-1, 1, len(pos))
    w = _params(9).astype(np.float64)
    gt, _ = _torch_grads(w, X, t)
    W1, b1, W2, b2, w3, b3 = M.split(w)
    f = M.forward(w, X, rounding=False)
    n = len(X)
    dz = 2 * (f["y"] - t) / n  # tanh derivative dropped
    wrong = f["H2"].T @ dz
    assert not np.allclose(wrong, gt[M.OFF_W3:M.OFF_B3], rtol=1e-3)
    dz = 2 * (f["y"] - t) * (1 - f["y"] ** 2) / n
    dH2 = np.where(f["H2"] > 0, dz[:, None] * w3[None, :], 0)
    dH1_wrong = np.where(f["H1"] > 0, dH2 @ W2, 0)  # W2 instead of W2^T
    wrong = (X.T @ dH1_wrong).ravel()
    assert not np.allclose(wrong, gt[M.OFF_W1:M.OFF_B1], rtol=1e-3)


def test_init_params_distribution(oracle):
    w = M.init_params(42, oracle.philox)
    W1, b1, W2, b2, w3, b3 = M.split(w)
    for W, fan in ((W1, 128), (W2, 256), (w3, 256)):
        a = np.sqrt(6 / fan)
        assert np.abs(W).max() <= a and np.abs(W).max() > 0.98 * a
        assert abs(W.mean()) < 0.05 * a
        assert abs(W.var() - a * a / 3) < 0.05 * a * a  # var of U(-a, a)
    assert not b1.any() and not b2.any() and not b3.any()


def test_epoch_order_is_a_permutation(oracle):
    p = M.epoch_order(1000, 7, 0, oracle.philox)
    assert sorted(p.tolist()) == list(range(1000))
    assert not np.array_equal(p, M.epoch_order(1000, 7, 1, oracle.philox))
