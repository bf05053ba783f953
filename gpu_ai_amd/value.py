"""Learned leaf value: data, training and error measurement on the device.

Positions are made as the playout workloads make them (seeded random plies from the initial
position, finished games rejected) and labelled by the method's own random playouts on the
device: target = (wins - losses) / playouts for the side to move, in [-1, 1].  The network
(include/gpuai/mi_mlp.h) is trained by the device's own forward, backward and Adam kernels.
Its error is measured against independent 1024-playout estimates of the same positions, next to
the error of k-playout estimates (k = 1 is what the reference's MCTS takes per leaf,
src/mcts.cpp:57-61), which prices one network evaluation in playouts.

    python -m gpu_ai_amd.value --positions 2000000 --epochs 24 --out value.ckpt
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from .engine import POLICY_RANDOM, Engine, ValueNet


def mover_value(pos: np.ndarray, tallies: np.ndarray) -> np.ndarray:
    """(wins - losses) / playouts for the side to move; tallies are (draws, p1 wins, p2 wins)."""
    t = tallies.astype(np.float64)
    n = t.sum(1)
    p1 = pos["turn"] == 0
    wins = np.where(p1, t[:, 1], t[:, 2])
    losses = np.where(p1, t[:, 2], t[:, 1])
    return ((wins - losses) / n).astype(np.float32)


def make_dataset(engine: Engine, n: int, seed: int, kmin: int = 0, kmax: int = 150,
                 playouts: int = 256, label_seed: int = 1, chunk: int = 1 << 18):
    """n positions (k uniform in [kmin, kmax] random plies, finished games rejected) and their
    playout labels."""
    pos = engine.start_positions(seed, n, kmin, kmax, reject_terminal=True)
    t = np.empty(n, np.float32)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        tallies, _ = engine.run_tally(pos[a:b], per_pos=playouts, seed=label_seed, epoch=a // chunk,
                                      policy=POLICY_RANDOM)
        t[a:b] = mover_value(pos[a:b], tallies)
    return pos, t


def train(net: ValueNet, pos: np.ndarray, t: np.ndarray, epochs: int, batch: int = 4096,
          lr: float = 1e-3, shuffle_seed: int = 7, log=None) -> list[dict]:
    net.set_dataset(pos, t)
    hist = []
    for ep in range(epochs):
        loss, ms = net.train_epoch(batch, lr, shuffle_seed, ep)
        steps = len(pos) // batch
        hist.append({"epoch": ep, "loss": loss, "device_ms": ms, "steps": steps,
                     "us_per_step": 1000.0 * ms / max(steps, 1)})
        if log:
            log(hist[-1])
    return hist


def evaluate(engine: Engine, net: ValueNet, n: int = 16384, seed: int = 99, kmin: int = 0,
             kmax: int = 90, ref_playouts: int = 1024, ks=(1, 4, 16, 64, 256)) -> dict:
    """Mean squared error against `ref_playouts`-playout estimates of held-out positions, for the
    network, a constant, and k-playout estimates drawn from independent streams."""
    pos = engine.start_positions(seed, n, kmin, kmax, reject_terminal=True)
    tallies, _ = engine.run_tally(pos, per_pos=ref_playouts, seed=seed + 1000)
    ref = mover_value(pos, tallies).astype(np.float64)
    y = np.concatenate([net.forward(pos[a:a + net.max_batch])
                        for a in range(0, n, net.max_batch)]).astype(np.float64)
    out = {"positions": n, "ref_playouts": ref_playouts,
           "mse_net": float(np.mean((y - ref) ** 2)),
           "mse_constant": float(np.mean((ref.mean() - ref) ** 2)),
           "mse_playouts": {}}
    for k in ks:
        tk, _ = engine.run_tally(pos, per_pos=k, seed=seed + 2000 + k)
        out["mse_playouts"][str(k)] = float(np.mean((mover_value(pos, tk) - ref) ** 2))
    # playouts per leaf that match the network's error (log-linear interpolation over ks)
    ks_arr = np.array(ks, dtype=np.float64)
    mse = np.array([out["mse_playouts"][str(k)] for k in ks])
    if out["mse_net"] >= mse[0]:
        eq = float(ks_arr[0] * mse[0] / out["mse_net"])
    elif out["mse_net"] <= mse[-1]:
        eq = float(ks_arr[-1])
    else:
        j = int(np.nonzero(mse <= out["mse_net"])[0][0])
        x0, x1 = np.log(ks_arr[j - 1]), np.log(ks_arr[j])
        y0, y1 = np.log(mse[j - 1]), np.log(mse[j])
        eq = float(np.exp(x0 + (np.log(out["mse_net"]) - y0) * (x1 - x0) / (y1 - y0)))
    out["equivalent_playouts"] = eq
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--positions", type=int, default=2_000_000)
    ap.add_argument("--playouts", type=int, default=256)
    ap.add_argument("--kmax", type=int, default=150,
                    help="training positions: k uniform in [0, kmax] random plies from the start")
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=2016)
    ap.add_argument("--eval", type=int, default=16384)
    ap.add_argument("--out", default="")
    ap.add_argument("--load", default="", help="skip training: load this checkpoint")
    ap.add_argument("--match-games", type=int, default=0,
                    help="then play this many games of network leaves against playout leaves")
    ap.add_argument("--move-ms", type=float, default=100.0)
    ap.add_argument("--leaf-batch", type=int, default=4096)
    ap.add_argument("--net-exploration", type=float, default=64.0,
                    help="k of the network player's UCB1 term sqrt(k ln N / n)")
    ap.add_argument("--playout-exploration", default="2",
                    help="k of the playout player's UCB1 term (the reference's is 2); a comma "
                         "list plays one match per value against the same network")
    ap.add_argument("--match-seed", type=int, default=2016, help="seed of the match openings")
    ap.add_argument("--mover-wins", action="store_true",
                    help="both trees count the chooser's own wins in UCB1 (the reference counts "
                         "the child's side to move)")
    args = ap.parse_args(argv)
    eng = Engine(0)
    with ValueNet(eng, args.batch, seed=args.seed) as net:
        if args.load:
            net.load(args.load)
        else:
            t0 = time.perf_counter()
            pos, t = make_dataset(eng, args.positions, args.seed, kmax=args.kmax,
                                  playouts=args.playouts)
            t_label = time.perf_counter() - t0
            print(json.dumps({"labelled": len(pos), "playouts": args.playouts, "kmax": args.kmax,
                              "seconds": t_label}), flush=True)
            train(net, pos, t, args.epochs, args.batch, args.lr,
                  log=lambda h: print(json.dumps(h), flush=True))
            del pos, t
        ev = evaluate(eng, net, args.eval)
        print(json.dumps({"eval": ev}), flush=True)
        if args.out:
            net.save(args.out)
        params = net.get_params()
    eng.close()
    if args.match_games:
        from . import host_api
        for kp in (float(x) for x in args.playout_exploration.split(",")):
            r = host_api.net_match(params, games=args.match_games, seed=args.match_seed,
                                   move_ms=args.move_ms, leaf_batch=args.leaf_batch,
                                   net_batch=args.batch, progress_every=25,
                                   net_exploration=args.net_exploration, playout_exploration=kp,
                                   mover_wins=args.mover_wins)
            r["outcome"] = "".join({1: "W", -1: "L", 0: "D"}[int(x)] for x in r["outcome"])
            r["net_exploration"] = args.net_exploration
            r["playout_exploration"] = kp
            r["mover_wins"] = args.mover_wins
            print(json.dumps({"match": r}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
