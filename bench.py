"""Throughput benchmark of the MI355X playout path (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]
                    [--no-match]

A step is one launch of the persistent playout kernel over the staged batch of the workload:
  random (default): 1024 seeded positions (k in [0, 30] random plies from the start) x 1024
                    random playouts = 1,048,576 bit-exact playouts
  heuristic       : 1024 seeded midgame positions (k in [20, 60], finished games redrawn) x 1024
                    heuristic playouts (largest weight, ties broken by the ply's word)
  noisy           : the same positions x 1024 playouts of the reference's heuristic, each move's
                    weight plus an N(0, 0.11) draw
  small           : 64 positions x 4096 random playouts (the host path's own size)
  multisearch     : 256 concurrent MCTS searches sharing the device (one line, its own fields)
  value           : value-network training and evaluation (one line, its own fields)
  cpu             : CPU rates only (reference host code and restatement, 1 and all threads)
Positions are generated on the device (mi_gen_start_positions) and staged in HBM once; each step
uses its own epoch (a fresh Philox stream), so no step repeats another.

roofline: the path is integer- and branch-bound.  Its work unit is one ply (generate, choose,
apply, terminal test) priced at a frozen W = 192 int32 lane-ops; achieved = plies x W / average
kernel time (HIP events on the library's stream), against the chip's int32 VALU issue rate
CUs x 64 lanes x clock; peak_measured is that rate as a probe kernel of plain v_add_u32 /
v_xor_b32 chains reaches it on this device (mi_valu_rate).  traffic: HBM bytes per launch from
rocprofv3 FETCH_SIZE/WRITE_SIZE runs recorded in profiles/ (null when absent); roofline.hbm sets
those bytes over the kernel time against the 8 TB/s HBM peak (the path is not memory bound).
cpu_baseline: the reference's own host code (oracle/_ref/libref.so: State::genMoves twice per ply
as src/playout.cpp:17-32 does, OpenMP over playouts, Philox draws) on a bounded sample of the same
workload, on the job's whole CPU share (OMP_NUM_THREADS, else every CPU it may use) and on one
thread, with the host's nproc, affinity and CPU model.
net_leaf_win_rate: the win-rate half of the metric, measured in the run (random workload, unless
--no-match or --tuning): the value network is trained on the device (8M positions labelled by 256
random playouts, 48 epochs, seconds), then 400 games of network leaves against playout leaves at
250 ms per move each (about 2.5 minutes; progress on stderr).  The network's score is
(wins + draws / 2) / games with its standard error.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "playouts/s on 1 MI355X; win rate of net-leaf vs playout MCTS at equal time"
W_PLY = 192                      # int32 lane-ops per ply (frozen cost model)
SPEC_CUS, SPEC_CLOCK_HZ = 256, 2.4e9
WORKLOADS = {
    "random": dict(n=1024, per_pos=1024, kmin=0, kmax=30, reject=False, policy=0,
                   name="random-1024x1024",
                   desc="random-1024x1024: 1024 seeded positions (k in [0,30] random plies) x "
                        "1024 bit-exact random playouts"),
    "heuristic": dict(n=1024, per_pos=1024, kmin=20, kmax=60, reject=True, policy=1,
                      name="heuristic-1024x1024",
                      desc="heuristic-1024x1024: 1024 seeded midgame positions (k in [20,60], "
                           "finished games redrawn) x 1024 heuristic playouts"),
    "noisy": dict(n=1024, per_pos=1024, kmin=20, kmax=60, reject=True, policy=2,
                  name="noisy-1024x1024",
                  desc="noisy-1024x1024: 1024 seeded midgame positions (k in [20,60], finished "
                       "games redrawn) x 1024 heuristic playouts with N(0, 0.11) weight noise"),
    "small": dict(n=64, per_pos=4096, kmin=0, kmax=30, reject=False, policy=0,
                  name="random-64x4096",
                  desc="random-64x4096: 64 seeded positions x 4096 random playouts"),
}
POS_SEED, PLAY_SEED = 2016, 12345


def host_call_overheads(eng) -> list:
    """Host-side time above the kernel of mi_run_playouts calls of MCTS batch sizes (host buffers
    in and out): the median over ten calls after two, per batch."""
    import numpy as np
    pos = eng.start_positions(POS_SEED + 1, 2048, 0, 30, False)
    batches = [("512 positions x 1", pos[:512], 1, None),
               ("2048 positions x 1-3 (4095 playouts, uneven counts as an MCTS batch)", pos,
                None, (np.arange(2048) % 3 + 1).astype(np.uint32)),
               ("1024 positions x 64", pos[:1024], 64, None)]
    out = []
    for name, p, per_pos, counts in batches:
        rows = []
        for e in range(12):
            st = eng.run_playouts(p, per_pos=per_pos, counts=counts, seed=PLAY_SEED,
                                  epoch=7000 + e).stats
            if e >= 2:
                rows.append((st["total_ms"], st["kernel_ms"]))
        out.append({"batch": name, "playouts": int(st["playouts"]),
                    "total_ms": round(statistics.median(r[0] for r in rows), 4),
                    "kernel_ms": round(statistics.median(r[1] for r in rows), 4),
                    "overhead_us": round(1e3 * statistics.median(r[0] - r[1] for r in rows), 1)})
    return out


def read_traffic(workload: str):
    """HBM bytes per launch recorded from rocprofv3 --pmc runs (profiles/*traffic*.md)."""
    path = os.path.join(ROOT, "profiles", f"k_playouts_{WORKLOADS[workload]['name']}_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def host_cpu() -> dict:
    """This host's processors: nproc, the CPUs this process may run on, the model name, and the
    thread share of the job (OMP_NUM_THREADS, which the GPU hosts set to one GPU's share)."""
    model = ""
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                     if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
    return {"nproc": os.cpu_count(), "affinity": affinity, "share": min(share, affinity),
            "model": model}


def cpu_baseline(cfg: dict, pos, threads: int) -> dict:
    """Times the reference build (or the C restatement when that is absent) on a bounded sample
    of the workload: on `threads` threads (the job's whole CPU share) and on one thread."""
    from oracle import pyoracle
    kind, lib = "reference", None
    if os.path.exists(pyoracle.REF_SO):
        lib = pyoracle.Reference(pyoracle.REF_SO)
    else:
        kind, lib = "port", pyoracle.Oracle()

    def timed(npos: int, per: int, th: int) -> tuple[float, str]:
        t0 = time.perf_counter()
        lib.run_playouts(pos[:npos], per_pos=per, seed=PLAY_SEED, epoch=0, policy=cfg["policy"],
                         threads=th, traces=False)
        dt = time.perf_counter() - t0
        return npos * per / dt, (f"first {npos} positions x {per} playouts of the workload "
                                 f"({npos * per} playouts, {dt:.2f} s wall on {th} threads)")

    # about 10-30 s of CPU work on all threads, a few seconds on one
    npos = min(len(pos), 1024 if cfg["policy"] == 0 else 256)
    per = min(cfg["per_pos"], 1024)
    rate, sample = timed(npos, per, threads)
    rate1, sample1 = timed(min(len(pos), 64), min(cfg["per_pos"], 1024), 1)
    return {"value": round(rate, 1), "unit": "playouts/s", "cores": threads, "kind": kind,
            "sample": sample, "one_thread": {"value": round(rate1, 1), "sample": sample1},
            "host": host_cpu()}


def bench_cpu(args) -> int:
    """CPU rates beside the device figure, on the random-64x4096 positions (the host path's own
    size).  Codes: the reference's own host code (oracle/_ref/libref.so: State::genMoves twice
    per ply as src/playout.cpp:17-32 does) with the Philox choice, and with glibc rand() % n as
    RandomPlayer::getMove draws it (src/player.cpp:13-16; rand() takes a global lock, so that
    variant shows the reference's contention); and the C restatement (oracle/liboracle.so).  Each
    on one thread and on all threads.  No GPU is used: the positions come from the restatement's
    start-position generator (the positions the device generates for the same seed)."""
    from oracle import pyoracle
    cfg = WORKLOADS["small"]
    orc = pyoracle.Oracle()
    pos = orc.start_positions(POS_SEED, cfg["n"], cfg["kmin"], cfg["kmax"])
    threads = args.cpu_threads or host_cpu()["share"]
    codes = [("restatement (oracle/liboracle.so)", orc, [("philox", 0)])]
    if os.path.exists(pyoracle.REF_SO):
        codes.insert(0, ("reference host code (oracle/_ref/libref.so)",
                         pyoracle.Reference(pyoracle.REF_SO),
                         [("philox", 0), ("glibc rand()", pyoracle.POLICY_RANDOM_LIBC)]))
    rows = []
    for code, lib, choices in codes:
        for choice, policy in choices:
            for th in sorted({1, threads}):
                # whole workload where it takes a second or two, else 512 playouts per position
                per = cfg["per_pos"] if th > 1 and policy == 0 else 512
                t0 = time.perf_counter()
                lib.run_playouts(pos, per_pos=per, seed=PLAY_SEED, epoch=0, policy=policy,
                                 threads=th, traces=False)
                dt = time.perf_counter() - t0
                rows.append({"code": code, "choice": choice, "threads": th,
                             "playouts": len(pos) * per, "seconds": round(dt, 3),
                             "playouts_per_s": round(len(pos) * per / dt, 1)})
                print(f"[cpu] {code}, {choice}, {th} threads: {rows[-1]['playouts_per_s']:.0f} "
                      f"playouts/s", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "CPU playouts/s (random-64x4096 positions)", "unit": "playouts/s",
                      "higher_is_better": True, "data": "synthetic (seeded positions)",
                      "host": host_cpu(), "rows": rows}), flush=True)
    return 0


def bench_multisearch() -> int:
    """Batched multi-search: 256 concurrent searches, budgets uniform in [1000, 64000] playouts,
    UCT c = sqrt(2).  One graph replay (persistent refill kernel) per iteration, the same with
    two halves of the searches overlapped with the host's tree work (pipelined), and one playout
    launch per search per iteration (the baseline); each mode's launch shape comes from the
    autotuner, run before its timed search."""
    from gpu_ai_amd import host_api
    runs = {m: host_api.multisearch(mode=m) for m in ("graph", "pipelined", "per_search")}
    g, pl, p = runs["graph"], runs["pipelined"], runs["per_search"]
    best = max(("graph", g), ("pipelined", pl), key=lambda kv: kv[1]["playouts_per_second"])
    # per-search arrays (budgets, root trials) stay out of the line; their totals go in
    fmt = lambda r: {**{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()
                        if not hasattr(v, "shape")},
                     "budget_total": int(r["budgets"].sum()),
                     "root_trials_total": int(r["root_trials"].sum())}
    line = {
        "metric": "aggregate playouts/s of 256 concurrent MCTS searches on 1 MI355X",
        "value": round(best[1]["playouts_per_second"], 1), "unit": "playouts/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "int32", "data": "synthetic (seeded positions)",
        "config": {"workload": "multisearch-256: 256 searches from seeded distinct positions, budgets "
                               "uniform in [1000, 64000] playouts, 256 trials per search "
                               "per iteration, UCT c=sqrt(2)", "mode": best[0],
                   "launch_shape": {k: best[1][k] for k in ("block_size", "blocks_per_cu",
                                                            "refill", "defer_min")},
                   "autotuned": True},
        "graph": fmt(g), "pipelined": fmt(pl), "per_search": fmt(p),
        "speedup_vs_per_search": round(best[1]["playouts_per_second"] / p["playouts_per_second"], 3),
    }
    print(json.dumps(line), flush=True)
    return 0


# The match: each player's exploration factor k in sqrt(k ln N / n), both at the best value of
# the sweeps at 250 ms per move (profiles/value_net_match_400x250ms.md): the strongest playout-only
# player over k in {0.5, 1, ..., 1024} (the reference's own player uses 2, src/mcts.cpp:189-190)
# and the network player's best k against it.
MATCH_NET_K = 1024.0
MATCH_PLAYOUT_K = 256.0
# The network of the match: value MSE 0.0055 against 1024-playout estimates (2M positions x 24
# epochs: 0.0088; the match score did not change measurably between them)
MATCH_POSITIONS, MATCH_EPOCHS = 8_000_000, 48


def measured_match(eng, args) -> dict:
    """Trains the value network in the run (8M positions labelled by 256 random playouts each,
    k in [0, 150] plies, 48 epochs of batch 4096, Adam lr 1e-3) and plays the 400-game match of
    network leaves against playout leaves at equal wall time per move (netmatch.hpp): the
    measured half of the metric.  Progress lines go to stderr."""
    from gpu_ai_amd import host_api, value
    from gpu_ai_amd.engine import ValueNet

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    pos, t = value.make_dataset(eng, MATCH_POSITIONS, POS_SEED, kmax=150, playouts=256)
    label_s = time.perf_counter() - t0
    log(f"labelled {len(pos)} positions with 256 playouts each in {label_s:.1f} s")
    with ValueNet(eng, 4096, seed=POS_SEED) as net:
        t1 = time.perf_counter()
        hist = value.train(net, pos, t, MATCH_EPOCHS, 4096, 1e-3)
        train_s = time.perf_counter() - t1
        del pos, t
        ev = value.evaluate(eng, net, 16384, ks=(1, 16, 256))
        params = net.get_params()
    log(f"trained {MATCH_EPOCHS} epochs in {train_s:.1f} s (loss {hist[-1]['loss']:.4f}, value MSE "
        f"{ev['mse_net']:.4f}); playing {args.match_games} games at {args.match_ms:g} ms per move")
    t2 = time.perf_counter()
    r = host_api.net_match(params, games=args.match_games, seed=POS_SEED, move_ms=args.match_ms,
                           progress_every=25, net_exploration=MATCH_NET_K,
                           playout_exploration=args.match_playout_k)
    match_s = time.perf_counter() - t2
    log(f"match: network {r['net_wins']} wins, playouts {r['playout_wins']} wins, "
        f"{r['draws']} draws, score {r['score']:.3f} in {match_s:.0f} s")
    return {"score": round(r["score"], 4), "stderr": round(r["score_stderr"], 4),
            "games": r["games"], "net_wins": r["net_wins"], "playout_wins": r["playout_wins"],
            "draws": r["draws"], "move_ms": args.match_ms, "openings_seed": POS_SEED,
            "net_exploration": MATCH_NET_K, "playout_exploration": args.match_playout_k,
            "leaves": {"network": r["net_leaves"], "playouts": r["playout_leaves"]},
            "training": {"positions": MATCH_POSITIONS, "labels": "256 random playouts",
                         "epochs": MATCH_EPOCHS, "label_s": round(label_s, 2),
                         "train_s": round(train_s, 2), "final_loss": round(hist[-1]["loss"], 5),
                         "value_mse": round(ev["mse_net"], 5),
                         "mse_playouts": {k: round(v, 5) for k, v in ev["mse_playouts"].items()}},
            "seconds": round(match_s, 1)}


# FLOPs of one training step of the 128-256-256-1 network on a batch of B positions: forward
# 2B(128*256 + 256*256 + 256), backward 2B(128*256 + 2*256*256) (dW1, dW2, dH1) + 2B*2*256
# (dw3, dH2): 262,912 multiply-adds per position
def mlp_step_flops(batch: int) -> int:
    fwd = 2 * batch * (128 * 256 + 256 * 256 + 256)
    bwd = 2 * batch * (128 * 256 + 2 * 256 * 256 + 2 * 256)
    return fwd + bwd


def bench_value(args) -> int:
    """Learned leaf value: label 2M positions with 256 random playouts each, train the network
    (batch 4096, Adam lr 1e-3) and measure its error against 1024-playout estimates."""
    from gpu_ai_amd import value
    from gpu_ai_amd.engine import Engine, ValueNet
    eng = Engine(0)
    t0 = time.perf_counter()
    pos, t = value.make_dataset(eng, 2_000_000, POS_SEED, playouts=256)
    label_s = time.perf_counter() - t0
    epochs = max(1, args.steps)
    with ValueNet(eng, 4096, seed=POS_SEED) as net:
        hist = value.train(net, pos, t, args.warmup, 4096, 1e-3) if args.warmup else []
        t1 = time.perf_counter()
        timed = value.train(net, pos, t, epochs, 4096, 1e-3, shuffle_seed=11)
        train_s = time.perf_counter() - t1
        ev = value.evaluate(eng, net, 16384)
    eng.close()
    steps = sum(h["steps"] for h in timed)
    device_s = sum(h["device_ms"] for h in timed) / 1e3
    flops = mlp_step_flops(4096)
    achieved = flops * steps / device_s
    line = {
        "metric": "value-network training throughput (positions/s) on 1 MI355X",
        "value": round(4096 * steps / train_s, 1), "unit": "positions/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(train_s / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (seeded positions labelled by device playouts)",
        "config": {"workload": "value-2Mx256: 2,000,000 positions labelled by 256 random "
                               "playouts, MLP 128-256-256-1, batch 4096, Adam lr 1e-3",
                   "epochs_timed": epochs, "epochs_warmup": args.warmup},
        "labelling": {"seconds": round(label_s, 3),
                      "playouts_per_s": round(2_000_000 * 256 / label_s, 1)},
        "value_error": {"mse_net": round(ev["mse_net"], 5),
                        "mse_playouts": {k: round(v, 5) for k, v in ev["mse_playouts"].items()},
                        "equivalent_playouts": round(ev["equivalent_playouts"], 1),
                        "after_epochs": epochs + args.warmup},
        "roofline": {"bound": "mfma", "achieved": round(achieved / 1e12, 3), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(achieved / 2.5e15, 5), "traffic": None,
                     "kernel": "training step (k_train_rows, k_wgrad_part, k_adam)",
                     "flops_per_step": flops, "device_us_per_step": round(device_s / steps * 1e6, 2)},
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="random",
                    choices=sorted(WORKLOADS) + ["multisearch", "value", "cpu"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch shape instead of tuning on the batch")
    ap.add_argument("--tuning", default="", metavar="BLOCK,BLOCKS_PER_CU,REFILL,DEFER",
                    help="use this launch shape instead of tuning (refill 0 static, 1 queue); "
                         "profiled runs use it so that only timed launches are traced")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run a plain copy kernel of BYTES (traffic-counter calibration)")
    ap.add_argument("--no-match", action="store_true",
                    help="skip the in-run training and network-vs-playout match (the random "
                         "workload runs them by default)")
    ap.add_argument("--match-games", type=int, default=400)
    ap.add_argument("--match-ms", type=float, default=250.0)
    ap.add_argument("--match-playout-k", type=float, default=MATCH_PLAYOUT_K)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the playout path runs on a single MI355X here "
              "(sharding over several GPUs is designed, not run)", file=sys.stderr)
        return 2

    if args.workload == "cpu":
        return bench_cpu(args)
    if args.workload == "multisearch":
        return bench_multisearch()
    if args.workload == "value":
        return bench_value(args)

    from gpu_ai_amd.engine import Engine

    cfg = WORKLOADS[args.workload]
    eng = Engine(0)
    info = eng.device_info()
    pos = eng.start_positions(POS_SEED, cfg["n"], cfg["kmin"], cfg["kmax"], cfg["reject"])
    tuned = None
    if args.tuning:
        b, bpc, refill, defer = (int(x) for x in args.tuning.split(","))
        eng.set_tuning(b, bpc, refill, defer)
    elif not args.no_autotune:
        # the autotuner times every candidate launch shape on this batch, five launches each (its
        # own Philox epochs), and keeps the one of smallest median time
        tuned, _ = eng.autotune(pos, per_pos=cfg["per_pos"], policy=cfg["policy"], reps=5)
    total = eng.stage(pos, per_pos=cfg["per_pos"])
    for i in range(args.warmup):
        eng.launch_staged(seed=PLAY_SEED, epoch=1000 + i, policy=cfg["policy"])

    # every launch returns after its HIP stream has synchronised, so the device is idle at both
    # ends of the timed region
    kernel_ms, plies = [], 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = eng.launch_staged(seed=PLAY_SEED, epoch=i, policy=cfg["policy"])
        kernel_ms.append(st["kernel_ms"])
        plies += st["plies"]
    elapsed = time.perf_counter() - t0
    tuning = eng.tuning

    value = total * args.steps / elapsed
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    plies_per_launch = plies / args.steps
    clock_hz = info["clock_khz"] * 1e3 if info["clock_khz"] else SPEC_CLOCK_HZ
    peak_dev = info["compute_units"] * 64 * clock_hz
    peak_spec = SPEC_CUS * 64 * SPEC_CLOCK_HZ
    achieved = plies_per_launch * W_PLY / avg_kernel_s
    traffic = read_traffic(args.workload)
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "playouts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded positions generated on the device)",
        "config": {"workload": cfg["desc"], "positions": cfg["n"],
                   "playouts_per_position": cfg["per_pos"], "playouts_per_step": total,
                   "policy": ("random", "heuristic", "heuristic_noisy")[cfg["policy"]],
                   "position_seed": POS_SEED, "playout_seed": PLAY_SEED,
                   "block_size": tuning["block_size"], "blocks_per_cu": tuning["blocks_per_cu"],
                   "refill": "queue" if tuning["refill"] == 1 else "static",
                   "defer_min": tuning["defer_min"], "autotuned": tuned is not None,
                   "longest_first": total >= 256 * info["compute_units"]},
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 4),
                     "peak": round(peak_spec / 1e12, 3), "unit": "Tops/s (int32 lane-ops)",
                     "frac": round(achieved / peak_spec, 4), "traffic": traffic,
                     "kernel": "k_playouts", "avg_kernel_ms": round(avg_kernel_s * 1e3, 4),
                     "plies_per_launch": int(plies_per_launch), "ops_per_ply": W_PLY,
                     "peak_device_clock": round(peak_dev / 1e12, 3),
                     "device": {"compute_units": info["compute_units"],
                                "clock_mhz": info["clock_khz"] / 1e3}},
    }
    # the same batch through mi_run_playouts with host buffers (positions in, winners out, over
    # PCIe): reported beside the figure, never as it.  Two calls set up the context's packed
    # batch buffers and HIP's copy paths; the medians of the next ten calls are reported.
    calls = []
    for e in range(12):
        st = eng.run_playouts(pos, per_pos=cfg["per_pos"], seed=PLAY_SEED, epoch=5000 + e,
                              policy=cfg["policy"]).stats
        if e >= 2:
            calls.append((st["total_ms"], st["kernel_ms"]))
    hb_total = statistics.median(c[0] for c in calls)
    line["host_buffers"] = {"total_ms": round(hb_total, 4),
                            "kernel_ms": round(statistics.median(c[1] for c in calls), 4),
                            "overhead_ms": round(statistics.median(c[0] - c[1] for c in calls), 4),
                            "calls": len(calls),
                            "playouts_per_s": round(total / (hb_total / 1e3), 1)}
    # MCTS-sized calls in autotuned runs only: a run with a fixed shape (as profiled runs are)
    # launches the 1M-playout batch alone
    if args.workload == "random" and not args.tuning:
        line["host_buffers"]["mcts_sized"] = host_call_overheads(eng)
    # the HBM roofline beside the VALU one: measured bytes per launch over the kernel time
    if traffic:
        line["roofline"]["hbm"] = {"achieved": round(traffic / avg_kernel_s / 1e9, 3),
                                   "peak": 8000.0, "unit": "GB/s",
                                   "frac": round(traffic / avg_kernel_s / 8e12, 6)}
    # the roofline's peak, measured: a probe kernel of independent int32 add/xor instructions
    valu_rate, valu_ms = eng.valu_rate()
    line["roofline"]["peak_measured"] = round(valu_rate / 1e12, 3)
    line["roofline"]["frac_of_measured"] = round(achieved / valu_rate, 4) if valu_rate else None
    if args.copy_calibration:
        ms = eng.copy_calibration(args.copy_calibration, 5)
        line["copy_calibration"] = {"bytes": args.copy_calibration, "avg_ms": round(ms, 4),
                                    "GB/s": round(2 * args.copy_calibration / ms / 1e6, 1)}
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_cpu()["share"]
        try:
            line["cpu_baseline"] = cpu_baseline(cfg, pos, threads)
        except Exception as e:  # the baseline is reported, never required for the GPU figure
            line["cpu_baseline"] = {"value": None, "error": str(e)}
    if args.workload == "random" and not args.no_match and not args.tuning:
        try:
            line["net_leaf_win_rate"] = measured_match(eng, args)
        except Exception as e:  # the throughput line is printed whatever happens to the match
            print(f"[bench] match failed: {e}", file=sys.stderr, flush=True)
            line["net_leaf_win_rate"] = {"score": None, "error": str(e)}
    else:
        line["net_leaf_win_rate"] = None  # skipped: --no-match, a profiled run, or another workload
    eng.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
