"""CPU restatement of the value network (TEST INFRASTRUCTURE ONLY).

The 128-256-256-1 MLP of gpu_ai_amd's value network (include/gpuai/mi_mlp.h), restated in numpy
for the parity tests of the device kernels: the input encoding, the forward pass with the same
BF16 roundings (inputs, weights, stored activations, round to nearest even), the gradient of the
mean squared error and the Adam step.  The reference has no learned component (its leaves are
played out, src/mcts.cpp:57-61), so there is no reference code to follow; each function is pinned
instead by tests/test_mlp_oracle.py against what fixes it independently: torch's bfloat16
conversion, torch autograd on the same network in float64 (rounding disabled), central finite
differences, torch.optim.Adam, and closed forms of the encoding.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 128, 256
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 32768, 33024, 98560, 98816, 99072
N_PARAMS = 99073


def bf16(x: np.ndarray) -> np.ndarray:
    """Rounds float32 values to BF16 (nearest, ties to even) and returns them as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _brev32(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint32)
    out = np.zeros_like(x)
    for b in range(32):
        out |= ((x >> np.uint32(b)) & np.uint32(1)) << np.uint32(31 - b)
    return out


def encode(pos: np.ndarray) -> np.ndarray:
    """Positions (POS_DTYPE) -> (n, 128) float32 planes in the frame of the side to move:
    own men, own kings, opponent men, opponent kings; player 2 sees the board rotated by 180
    degrees (square s -> 31 - s)."""
    p1 = pos["p1"].astype(np.uint32)
    p2 = pos["p2"].astype(np.uint32)
    k = pos["kings"].astype(np.uint32)
    t = pos["turn"].astype(bool)
    own = np.where(t, _brev32(p2), p1)
    opp = np.where(t, _brev32(p1), p2)
    kings = np.where(t, _brev32(k), k) & (own | opp)
    planes = [own & ~kings, own & kings, opp & ~kings, opp & kings]
    X = np.zeros((len(pos), N_IN), dtype=np.float32)
    for j, bits in enumerate(planes):
        for s in range(32):
            X[:, 32 * j + s] = (bits >> np.uint32(s)) & np.uint32(1)
    return X


def mirror(pos: np.ndarray) -> np.ndarray:
    """The same positions with the colours exchanged and the board rotated by 180 degrees."""
    m = pos.copy()
    m["p1"], m["p2"] = _brev32(pos["p2"]), _brev32(pos["p1"])
    m["kings"] = _brev32(pos["kings"])
    m["turn"] = 1 - pos["turn"]
    return m


def split(w: np.ndarray):
    """Views of the parameter vector: W1 (128,256), b1, W2 (256,256), b2, w3 (256,), b3 (1,)."""
    return (w[OFF_W1:OFF_B1].reshape(N_IN, N_HID), w[OFF_B1:OFF_W2],
            w[OFF_W2:OFF_B2].reshape(N_HID, N_HID), w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3],
            w[OFF_B3:])


def forward(w: np.ndarray, X: np.ndarray, rounding: bool = True) -> dict:
    """Forward pass.  With rounding, operands and stored activations are BF16 as on the device
    (sums in float64, cast to float32); without, everything is float64."""
    W1, b1, W2, b2, w3, b3 = split(w)
    r = bf16 if rounding else (lambda a: a)
    ft = np.float32 if rounding else np.float64
    Xr = X.astype(ft)
    z1 = (Xr.astype(np.float64) @ r(W1).astype(np.float64)).astype(ft) + b1.astype(ft)
    H1 = r(np.maximum(z1, 0).astype(ft))
    z2 = (H1.astype(np.float64) @ r(W2).astype(np.float64)).astype(ft) + b2.astype(ft)
    H2 = r(np.maximum(z2, 0).astype(ft))
    z3 = (H2.astype(np.float64) @ w3.astype(np.float64)).astype(ft) + b3.astype(ft)
    y = np.tanh(z3)
    return {"X": Xr, "H1": H1, "H2": H2, "y": y, "rounding": rounding}


def gradients(w: np.ndarray, X: np.ndarray, targets: np.ndarray, rounding: bool = True):
    """Gradient of (1/n) sum (y - t)^2 with respect to every parameter (parameter layout) and the
    loss.  With rounding, the back-propagated activation gradients dH2, dH1 are BF16 like the
    device's MFMA operands."""
    W1, b1, W2, b2, w3, b3 = split(w)
    f = forward(w, X, rounding)
    r = bf16 if rounding else (lambda a: a)
    ft = np.float32 if rounding else np.float64
    n = len(X)
    y, H1, H2 = f["y"], f["H1"], f["H2"]
    t = targets.astype(ft)
    e = y - t
    loss = float(np.sum(e.astype(np.float64) ** 2) / n)
    dz = (2 * e * (1 - y * y) / n).astype(ft)
    g = np.zeros(N_PARAMS, dtype=np.float64)
    g[OFF_W3:OFF_B3] = H2.astype(np.float64).T @ dz.astype(np.float64)
    g[OFF_B3] = dz.astype(np.float64).sum()
    dH2 = r(np.where(H2 > 0, dz[:, None] * w3.astype(ft)[None, :], 0).astype(ft))
    g[OFF_B2:OFF_W3] = dH2.astype(np.float64).sum(0)
    g[OFF_W2:OFF_B2] = (H1.astype(np.float64).T @ dH2.astype(np.float64)).ravel()
    back = (dH2.astype(np.float64) @ r(W2).astype(np.float64).T).astype(ft)
    dH1 = r(np.where(H1 > 0, back, 0).astype(ft))
    g[OFF_B1:OFF_W2] = dH1.astype(np.float64).sum(0)
    g[OFF_W1:OFF_B1] = (f["X"].astype(np.float64).T @ dH1.astype(np.float64)).ravel()
    return g, loss


def loss_only(w: np.ndarray, X: np.ndarray, targets: np.ndarray) -> float:
    """Float64 loss without rounding (for finite differences)."""
    y = forward(w.astype(np.float64), X, rounding=False)["y"]
    return float(np.mean((y - targets.astype(np.float64)) ** 2))


class Adam:
    """Adam (Kingma & Ba, 2015, Algorithm 1) with bias correction, in float32."""

    def __init__(self, n: int, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
                 eps: float = 1e-8):
        self.m = np.zeros(n, np.float32)
        self.v = np.zeros(n, np.float32)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, beta1, beta2, eps

    def step(self, w: np.ndarray, g: np.ndarray) -> np.ndarray:
        """One update; the scalar coefficients (1 - beta, bias corrections) are formed in
        float64 and rounded to float32 once, the moments are float32."""
        f = np.float32
        self.t += 1
        g = g.astype(f)
        self.m = f(self.b1) * self.m + f(1 - self.b1) * g
        self.v = f(self.b2) * self.v + f(1 - self.b2) * g * g
        mhat = self.m * f(1 / (1 - self.b1 ** self.t))
        vhat = self.v * f(1 / (1 - self.b2 ** self.t))
        return (w - f(self.lr) * mhat / (np.sqrt(vhat) + f(self.eps))).astype(f)


def init_params(seed: int, philox) -> np.ndarray:
    """He-uniform initial parameters: U(-a, a), a = sqrt(6 / fan_in), from Philox key (seed,
    tensor), counter (i >> 2, 0, 3, 0), word i & 3 mapped to [0, 1) by its top 24 bits; biases 0.
    `philox(ctr, key)` returns the four words of one block."""
    w = np.zeros(N_PARAMS, np.float32)
    for t, (off, count, fan_in) in enumerate(((OFF_W1, 128 * 256, 128), (OFF_W2, 256 * 256, 256),
                                              (OFF_W3, 256, 256))):
        a = np.float32(np.sqrt(np.float32(6.0) / np.float32(fan_in)))
        words = np.concatenate([philox([i, 0, 3, 0], [seed, t]) for i in range((count + 3) // 4)])
        u = (words[:count] >> 8).astype(np.float32) * np.float32(1.0 / 16777216.0)
        w[off:off + count] = (np.float32(2.0) * u - np.float32(1.0)) * a
    return w


def epoch_order(n: int, shuffle_seed: int, epoch: int, philox) -> np.ndarray:
    """The training order of one epoch: Fisher-Yates from the last index down, the c-th swap
    drawing word c & 3 of Philox block (c >> 2, 0, 3, 1) under key (shuffle_seed, epoch)."""
    perm = np.arange(n, dtype=np.uint32)
    words = None
    for c, i in enumerate(range(n - 1, 0, -1)):
        if c & 3 == 0:
            words = philox([c >> 2, 0, 3, 1], [shuffle_seed, epoch])
        j = int(words[c & 3]) % (i + 1)
        perm[i], perm[j] = perm[j], perm[i]
    return perm
